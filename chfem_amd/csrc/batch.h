// batch.h -- tables and launchers of the batched solver (batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace chfem {
namespace batch {

// Walsh tile of the batched operator (rows x columns per block)
#define TILE_R 32
#define TILE_C 8

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
    SB_STRIDE = 16
};

enum { OP_ITER = 0, OP_REPL = 1 };
enum { FIN_B_ALPHA = 0, FIN_B_DELTA, FIN_B_TRUE, FIN_B_DELTA0, FIN_B_START };

struct Batch {
    int nsys = 0, nd = 1;
    const Sys *sys = nullptr;
    const Tile *tiles = nullptr;
    int ntiles = 0;
    const Chunk *chunks = nullptr;
    int nchunks = 0;
    double *Sb = nullptr;      // nsys * SB_STRIDE
    double *part = nullptr;    // max(ntiles, nchunks)
};

template <typename T> void apply(const Batch &b, const T *v, T *q, int mode, hipStream_t s);
void finish(const Batch &b, int what, hipStream_t s);
void dir(const Batch &b, float *d, const float *r, hipStream_t s);
void update(const Batch &b, float *x, float *r, const float *d, const float *q, hipStream_t s);
void start(const Batch &b, const double *r64, float *r32, float *d32, float *x32, hipStream_t s);
void delta0(const Batch &b, const double *r64, hipStream_t s);
void flush(const Batch &b, double *x64, double *e64, float *x32, hipStream_t s);
void resid(const Batch &b, double *r64, const double *q64, float *r32, hipStream_t s);

}  // namespace batch
}  // namespace chfem
