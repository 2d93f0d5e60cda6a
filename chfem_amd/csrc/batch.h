// batch.h -- tables and launchers of the batched solver (batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace chfem {
namespace batch {

// Walsh tile of the batched operators: 64 x 7 work-items own 64 node rows
// (the halo-wavefront variant of walsh_tile: one more wavefront evaluates
// the row above the tile), so 256 rows take 4 tiles; 512 work-items per block
#define TILE_R 64
#define TILE_C 7
#define TILE_HALO true
#define TILE_ROWS (TILE_HALO ? TILE_R : TILE_R - 1)
#define TILE_THREADS (TILE_R * TILE_C + (TILE_HALO ? 64 : 0))
// node layers per block
#define TILE_LZ 32

// one linear system (a load case of a sample), device view
struct Sys {
    int ny, nx, nz, LZ;       // grid (elements = nodes) and node layers per tile
    uint32_t N;               // nodes
    size_t off;               // DOF offset in the concatenated vectors
    const uint16_t *keys;     // packed NBN keys of the sample
    const void *mat32, *mat64;   // elastic: {lambda, mu} per material; thermal: {k, 0} per material
    const void *kinv32, *kinv64; // inverse Jacobi diagonal per key, [65536][nd]
    int tb0, tb1;             // operator tiles [tb0, tb1) of this system
    int vb0, vb1;             // vector chunks [vb0, vb1)
};

struct Tile {
    int sys, bx, by, bz;      // Walsh tile of system sys
    int slot;                 // its index in the system-contiguous order [tb0, tb1): its partials' position
};

struct Chunk {
    int sys;
    uint32_t i0, i1;          // node range, at most 256 nodes
};

// per-system FP64 scalars, row of SB_STRIDE doubles
enum {
    SB_DELTA = 0, SB_DELTA_OLD, SB_DELTA0, SB_DQ, SB_ALPHA, SB_THRESH, SB_ITER,
    SB_DONE,    // FP32 segment reached its threshold (set on the device)
    SB_ACTIVE,  // system still iterating
    SB_FRESH,   // next iteration is the first of a segment: direction = dinv r
    SB_REPL,    // residual replacement selected by the host
    SB_START,   // fresh FP32 start selected by the host
    SB_BPREV,   // single-kernel step: beta that formed the newest direction
    SB_BASE,    // single-kernel step: iteration count at the segment's start
    // alpha of each step of the current ring window (fused loop), zeroed
    // after every window flush
    SB_AHIST = 16, SB_RING_MAX = 16,
    SB_STRIDE = 32
};

enum { OP_ITER = 0, OP_REPL = 1 };
enum { FIN_B_ALPHA = 0, FIN_B_DELTA, FIN_B_TRUE, FIN_B_DELTA0, FIN_B_START, FIN_B_CG1 };

struct Batch {
    int nsys = 0, nd = 1;
    const Sys *sys = nullptr;
    const Tile *tiles = nullptr;
    int ntiles = 0;
    const Chunk *chunks = nullptr;
    int nchunks = 0;
    double *Sb = nullptr;      // nsys * SB_STRIDE
    double *part = nullptr;    // max(4 * ntiles, nchunks)
};

template <typename T> void apply(const Batch &b, const T *v, T *q, int mode, hipStream_t s);
void finish(const Batch &b, int what, hipStream_t s);
void dir(const Batch &b, float *d, const float *r, hipStream_t s);
void update(const Batch &b, float *x, float *r, const float *d, const float *q, hipStream_t s);
void start(const Batch &b, const double *r64, float *r32, float *d32, float *x32, hipStream_t s);
void delta0(const Batch &b, const double *r64, hipStream_t s);
void flush(const Batch &b, double *x64, double *e64, float *x32, hipStream_t s);
void resid(const Batch &b, double *r64, const double *q64, float *r32, hipStream_t s);

// Fused loop on z = M^-1 r (FP32): directions in a ring of w <= SB_RING_MAX
// slots of `total` floats each, indexed by each system's own iteration count
// j (SB_ITER).  A step: d = z + beta d_old (d = z on a system's fresh step;
// d_old in slot (j - 1) % w) into slot j % w, q = A d and partial d.q
// (apply_dir); alpha into SB_AHIST + j % w (finish_alpha); z -= alpha M^-1 q
// and partial z.(D z) (zupdate).  Every w steps ring_flush adds
// sum alpha_k d_k to x32 and zeroes the alphas (slots a system did not fill
// in the window have alpha 0).
void apply_dir(const Batch &b, const float *z, float *ring, size_t total, int w, float *q, hipStream_t s);
void finish_alpha(const Batch &b, int w, hipStream_t s);
void zupdate(const Batch &b, float *z, const float *q, hipStream_t s);
void ring_flush(const Batch &b, float *x32, const float *ring, size_t total, int w, hipStream_t s);
// z-form start and replacement: as start / resid with z32 = M^-1 r in place
// of r32 (and no d32).  z2 non-null (single-kernel steps): a system's z goes
// to the buffer its next step reads, z32 when its iteration count j is even
// and z2 when it is odd.
void start_z(const Batch &b, const double *r64, float *z32, float *x32, hipStream_t s, float *z2 = nullptr);
void resid_z(const Batch &b, double *r64, const double *q64, float *z32, hipStream_t s, float *z2 = nullptr);
// Single-kernel step (the single-system loop's form 2 per system): the
// deferred z update z = z_old - alpha M^-1 q_old, d = z + beta d_old into
// ring slot j % w, q = A d stored as M^-1 q, and the partials of d.q, z.q,
// q.M^-1 q and z.D z, then one finishing block per system (FIN_B_CG1:
// alpha, the next r.z by the recurrence restarted from z.D z, the iteration
// counted).  z and q alternate between two buffers by the parity of j (step
// j reads buffer (j + 1) % 2 and writes buffer j % 2 of {z2, z} and {q,
// q2}); from a segment's third step on (SB_BASE + 2, with w >= 3) z is
// recovered from the two previous directions and neither read nor written.
void apply_cg1(const Batch &b, float *z, float *z2, float *q, float *q2, float *ring, size_t total, int w,
               hipStream_t s);

}  // namespace batch
}  // namespace chfem
