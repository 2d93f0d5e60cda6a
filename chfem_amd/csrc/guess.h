// guess.h -- launchers of the learned-initial-guess kernels (guess.hip).
//
// Tensors are channels-last over a periodic voxel grid: t[v * C + c] with the
// solver's voxel order v = row + col*ny + layer*nx*ny and C a multiple of 16.
// Activations and matrix-core operands are BF16 (stored as uint16_t);
// weights are kept in FP32 (master) and packed to BF16 per launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace chfem {
namespace nn {

struct Grid3 {
    int ny, nx, nz;
};

enum ConvMode {
    CONV_SAME = 0,    // 3x3x3, stride 1, periodic
    CONV_DOWN = 1,    // 2x2x2, stride 2 (output grid halves)
    CONV_UP = 2,      // 3x3x3 on the nearest-neighbour x2 upsampled input (output grid doubles)
    CONV_PARITY = 3,  // input gradient of CONV_DOWN: one tap per output parity class
};

// out = act(conv(in, W) + bias + skip).  Ci, Co multiples of 16 (Co <= 64).
// W is packed BF16 [Co][taps * Ci] (CONV_PARITY: [8][Co][Ci]).  out is BF16,
// or FP32 when out_f32 (then accumulate adds into it).  bias/skip nullable.
void conv(int mode, const uint16_t *in, Grid3 gi, int Ci, const uint16_t *W, const float *bias,
          const uint16_t *skip, int relu, void *out, int out_f32, int accumulate, Grid3 go, int Co, hipStream_t s);

// grad[co][t*Ci + ci] (+)= sum over output voxels v of dz[v][co] * in[u(v, t)][ci]
// for the forward conv of this mode; part is scratch of wpart_size() floats.
size_t wpart_size(Grid3 go, int Co, int Ci, int taps);
void conv_wgrad(int mode, const uint16_t *in, Grid3 gi, int Ci, const uint16_t *dz, Grid3 go, int Co,
                float *part, float *grad, hipStream_t s);

// FP32 master [Co][taps*Ci] -> BF16: how = 0 same layout, 1 flipped and
// transposed [Ci][taps*Co] (input gradient of SAME/UP), 2 per-tap transposed
// [taps][Ci][Co] (input gradient of DOWN)
void pack_weights(const float *Wm, uint16_t *out, int Co, int Ci, int taps, int how, hipStream_t s);

// dz = bf16(dy * (y > 0)) (relu: y is the BF16 post-activation output; without
// relu dz = bf16(dy)); skip_grad (nullable) += dy*(y>0); bias_grad[c] += sum_v dz
size_t bias_part_size(size_t V, int C);
void act_backward(const float *dy, const uint16_t *y, int relu, uint16_t *dz, float *skip_grad, float *bias_part,
                  float *bias_grad, size_t V, int C, hipStream_t s);

// coarse[w][c] (+)= sum of fine[child][c] over the 8 children of w
void pool_children(const float *fine, Grid3 gf, float *coarse, int C, int accumulate, hipStream_t s);

// dy[v][c] = 2 (y - target) * scale for c < nt (else 0); loss += scale * sum (y - target)^2
size_t loss_part_size(size_t V);
void mse(const float *y, int Cp, const float *target, int nt, size_t V, double scale, float *dy, double *part,
         double *loss, hipStream_t s);

void adam(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
          int step, float gscale, hipStream_t s);

// input features: in[v][0..nf) = table[elem[v] * nf + k], other channels 0
void features(const uint16_t *elem, const float *table, int nf, uint16_t *in, int C, size_t V, hipStream_t s);

// x0[nd*i + a] = ramp of load case `load` at node i, component a, plus
// scale * y[i][load*nd + a]
void guess_to_x0(const float *y, int Cp, int load, int nd, Grid3 g, float scale, double *x0, hipStream_t s);
// the same for every load case with x0[load] != nullptr (6 elastic, 3
// thermal), in one pass over y: y holds each node's outputs of all load
// cases together, so one pass reads it once instead of once per load case
void guess_to_x0_all(const float *y, int Cp, int nd, Grid3 g, float scale, double *const x0[6], hipStream_t s);

// fill / convert helpers
void fill_f32(float *p, float v, size_t n, hipStream_t s);
void add_f32(float *dst, const float *src, size_t n, hipStream_t s);

// lane-layout check of the BF16 matrix-core product: D = A (16x16) * B (16x16)
void mfma_check(const uint16_t *A, const uint16_t *B, float *D, hipStream_t s);

}  // namespace nn

// network set by hmgSetInitialGuessNetwork (nullptr: none)
extern void *guess_active;
// prediction of every load case's fluctuation field for an element map
// (host); returns the device tensor y[V][*Cout] (valid until the next call
// on this network, ready once guess_stream(net) is synchronised) or nullptr
// when the grid is not a multiple of 8 in every direction
const float *guess_predict_device(void *net, const uint16_t *elem_host, const double *props, unsigned nmat,
                                  unsigned nx, unsigned ny, unsigned nz, int *Cout);
hipStream_t guess_stream(void *net);
// physics of a network (0 thermal, 1 elastic), -1 for nullptr
int guess_physics(void *net);

}  // namespace chfem
