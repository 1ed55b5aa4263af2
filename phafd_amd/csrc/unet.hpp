// unet.hpp -- the learned surrogate: a three-level 3D convolutional U-Net on
// the periodic grid (16 / 32 / 64 channels, 3^3 kernels), BF16 weights and
// activations, FP32 accumulation on the MFMA units, hand-written forward and
// backward passes and an Adam trainer.
//
// Input: 8 channels per grid point (4 used: the deviation of phi from its
// mean, the deposited bead density at the start and at the end of the
// block, the mean of phi; 4 zero); output: the change of phi over K steps
// beyond the exponential-Euler base, scaled.  Layers (L1 = the grid, L2 and L3 its
// 2x and 4x coarsenings by 2^3 averaging):
//   0 e1   conv3  8->16 relu  L1      5 d2   conv3 32->32 relu  L2 (input up(q2) + s2)
//   1 e1b  conv3 16->16 relu  L1 (s1) 6 q1   conv1 32->16       L2
//   2 e2   conv3 16->32 relu  L2 (s2) 7 d1   conv3 16->16 relu  L1 (input up(q1) + s1)
//   3 b    conv3 32->64 relu  L3      8 head conv1 16->1        L1 (FP32 out)
//   4 q2   conv1 64->32       L3
// Parameters at the boundary are in PyTorch Conv3d order per layer (weight
// [cout][cin][k][k][k], then bias), layer 0 with UNET_IN input channels.
#pragma once
#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace phafd {

constexpr int UNET_IN = 4;   // input channels used (of the 8 stored)

struct UNetLayer {
  int cin, cout, taps, level;   // cin as stored (layer 0: 8)
  bool relu;
  size_t w, b;                  // offsets of weight / bias in the parameter vector
  size_t pf, pt;                // offsets of the packed forward / transposed weights (bf16)
};

struct UNetActs {   // activations and their gradients for one grid shape
  int nx = 0, ny = 0, nz = 0;
  bool train = false;
  DBuf<uint16_t> ae1, s1, p1, s2, p2, a3, q2, t2, a2, q1, t1, a1;
  DBuf<uint16_t> g1a, g1b, g1c, g2a, g2b, g2c, g3a, g3b;   // gradients (bf16)
  DBuf<float> out, dout, part;
  DBuf<double> red;
};

struct UNet {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<UNetLayer> L;
  size_t nparam = 0, npack = 0;
  DBuf<float> P, G, Mo, Vo;              // FP32 master parameters, gradients, Adam moments
  DBuf<uint16_t> Wp;                     // packed bf16 weights (forward and transposed)
  bool packed = false;
  int64_t adam_t = 0;
  double norm[5] = {0.02, 1.0, 0.4, 0.1, 0.01};   // phi scale, rho scale, mean ref, mean scale, output scale
  UNetActs A;
  // training set: inputs [S][N][8] bf16, targets [S][N] float
  DBuf<uint16_t> dsX;
  DBuf<float> dsY;
  int ds_n = 0, ds_nx = 0, ds_ny = 0, ds_nz = 0;
  uint64_t rng = 0x9e3779b97f4a7c15ull;
  std::string err;

  void init(int device, uint64_t seed);
  void pack(hipStream_t s);
  void ensure(int nx, int ny, int nz, bool train);
  // out[N] (FP32, scaled change of phi) from X0[N][8] (bf16)
  void forward(const uint16_t *X0, float *out, hipStream_t s);
  // gradients of sum_v dout[v] * out[v] added to G (after forward on the same input)
  void backward(const uint16_t *X0, const float *dout, hipStream_t s);
  // mean square error of out against target, dout = its gradient times `scale`
  double loss_grad(const float *target, double scale, hipStream_t s);
  void adam(double lr, double b1, double b2, double eps, hipStream_t s);
};

// grid-shape requirement of the network: nx % 64 == 0, ny % 32 == 0, nz % 4 == 0
bool unet_shape_ok(int nx, int ny, int nz);
// X0 (bf16 [N][8]) from phi (FP64), the bead densities at the block's start
// and end (FP32) and the FP64 sum of phi; norm as in UNet::norm
void unet_features(size_t N, const double *phi, const float *rho0, const float *rho1, const double *phisum,
                   const double *norm, uint16_t *X0, hipStream_t s);
// X0 from host-layout float channels [4][N]
void unet_features_from_channels(size_t N, const float *ch4, uint16_t *X0, hipStream_t s);
// phi_out = base + scale * (out - mean(out)); outsum: FP64 sum of out
void unet_apply(size_t N, const double *base, const float *out, const double *outsum, double scale, double *phi_out,
                hipStream_t s);

}  // namespace phafd

// the handle of the C ABI (phafd_hip.h)
struct phafd_unet {
  phafd::UNet net;
  phafd::DBuf<float> hin;           // staging of host arrays
  phafd::DBuf<uint16_t> x0;         // input of the last forward
};
