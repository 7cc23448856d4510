// lh_nn.h -- learned initial guess for the pressure solve: a 3D convolutional
// encoder-decoder attached to a scene context (k_nn.hip).
//
// Network (channels-last BF16 tensors, FP32 accumulation, 3x3x3 kernels, zero padding):
//   level 0 (N):    E0 conv 2->16  relu            D0 conv 48->16 relu  -> OUT conv 16->1
//   level 1 (N/2):  E1 conv 16->32 relu            D1 conv 96->32 relu
//   level 2 (N/4):  E2 conv 32->64 relu            D2 conv 160->64 relu
//   level 3 (N/8):  E3 conv 64->96 relu
// Down: 2x2x2 average pool.  Up: nearest-neighbour upsample, concatenated
// [upsampled, skip] before each decoder conv.
// Input channels: divergence / rms(divergence over liquid cells), cell type
// (+1 liquid, 0 air, -1 solid).  Output: the pressure on interior liquid cells up to a
// factor (trained against p / rms(p) / target_scale; the solver rescales the guess).
#pragma once
#include "lh_ctx.h"

namespace lh {

constexpr int kNnLayers = 8;

struct Net {
  // parameters: per layer W[cout][kp] (K index = tap * cin + ci, tap = (dx+1)*9 + (dy+1)*3 + (dz+1),
  // zero padded to kp = 32-multiple) then bias[cout]; FP32 master copy, BF16 copy, Adam moments
  size_t off[kNnLayers] = {}, boff[kNnLayers] = {}, wtoff[kNnLayers] = {};
  int kp[kNnLayers] = {}, kpt[kNnLayers] = {};
  size_t npad = 0, nlog = 0, nwt = 0;
  float *w32 = nullptr, *g32 = nullptr, *am = nullptr, *av = nullptr;
  uint16_t *w16 = nullptr, *wt16 = nullptr;  // BF16 weights; transposed, tap-flipped copies for backward data
  int adam_t = 0;
  float target_scale = 1.f;
  // training loss on the simulator's systems: 0 masked MSE against p / rms(p);
  // 1 energy: the A-norm error of the energy-scaled guess, 1 - (y.b)^2 / ((y.Ay)(x*.b))
  int loss_mode = 0;
  // activations for one grid of dims (all divisible by 8)
  int dims[3] = {0, 0, 0};
  bool train_bufs = false;
  uint16_t *x0 = nullptr, *a0 = nullptr, *p0 = nullptr, *a1 = nullptr, *p1 = nullptr, *a2 = nullptr, *p2 = nullptr,
           *a3 = nullptr, *c2 = nullptr, *d2 = nullptr, *c1 = nullptr, *d1 = nullptr, *c0 = nullptr, *d0 = nullptr;
  float *y = nullptr, *t = nullptr, *m = nullptr, *gy = nullptr, *warm = nullptr, *stage = nullptr;
  uint16_t *g_d0 = nullptr, *g_c0 = nullptr, *g_d1 = nullptr, *g_c1 = nullptr, *g_d2 = nullptr, *g_c2 = nullptr,
           *g_a3 = nullptr, *g_p2 = nullptr, *g_a2 = nullptr, *g_p1 = nullptr, *g_a1 = nullptr, *g_p0 = nullptr,
           *g_a0 = nullptr;
  double *part = nullptr, *sc = nullptr;  // block partials; scalars [0] 1/rms(b) [1] n liquid [2] n mask [3] loss [4] target factor
};

// channel counts and pyramid level of each layer (E0 E1 E2 E3 D2 D1 D0 OUT)
extern const int kNnCin[kNnLayers], kNnCout[kNnLayers], kNnLevel[kNnLayers];

int nn_create(Ctx& c, uint64_t seed);             // fresh He-initialised network on c.net
void nn_free(Net* n);
int nn_save(Ctx& c, const char* path);
int nn_load(Ctx& c, const char* path);
int nn_get_params(Ctx& c, float* host, int64_t count, bool grads);
int nn_set_params(Ctx& c, const float* host, int64_t count);
// forward on a host input (V x 2, channels last) of the given dims; y (V) to host
int nn_forward_host(Ctx& c, const int dims[3], const float* in, float* y);
// loss and gradient (logical layout) for host input/target/mask; optional Adam step with lr > 0
int nn_step_host(Ctx& c, const int dims[3], const float* in, const float* target, const float* mask, float lr,
                 float* grad, float* loss);
// loss, gradient and (lr > 0) one Adam step on the last built system and its solved pressure
int nn_train_last_solve(Ctx& c, float lr, float* loss);
// the last solve's training pair to host (input V x 2 as BF16-rounded floats, target V, mask V)
int nn_sample_last_solve(Ctx& c, float* in, float* target, float* mask);
// initial guess for A x = b on the built operator of c (b: device, level-0 layout): the
// network's prediction on interior liquid cells, scaled by the energy-minimising factor;
// *x0 = device buffer (level-0 layout)
int nn_predict(Ctx& c, const float* b, float** x0);
// x0 <- alpha x0 with alpha = (x0 . b) / (x0 . A x0) on the built level-0 operator (k_solver.hip)
void launch_warm_ritz(Ctx& c, float* x0, const float* b);
// ay = A y and out[0..2] = (y . b, y . A y, xs . b) on the built level-0 operator (k_solver.hip)
void launch_energy_terms(Ctx& c, const float* y, const float* b, const float* xs, float* ay, double* out);

}  // namespace lh
