// guess.cpp -- learned initial guess: a periodic 3D convolutional
// encoder-decoder that maps a voxel image's material properties to the
// fluctuation field of every load case, trained on solutions the library
// computes itself.  The solver seeds PCG with ramp + prediction and falls back
// to the zero guess when the prediction's initial residual is larger.
//
// Network (channels; grids g0 = image, g1..g3 halved each time):
//   e0 = relu(conv3(relu(conv3(features, 16)), 16))            g0
//   e1 = relu(conv3(relu(down2(e0, 32)), 32))                 g1
//   e2 = relu(conv3(relu(down2(e1, 48)), 48))                 g2
//   e3 = relu(conv3(relu(down2(e2, 48)), 48))                 g3
//   d2 = relu(up(e3, 48) + e2), d1 = relu(up(d2, 32) + e1), d0 = relu(up(d1, 16) + e0)
//   y  = conv3(d0, OUT)                                         g0, FP32
// conv3: 3x3x3 periodic; down2: 2x2x2 stride 2; up: 3x3x3 on the x2
// nearest-neighbour upsampled input.  OUT = 16 (thermal: 3 used) or 32
// (elastic: 18 used = 6 load cases x 3 components).  Parameters (FP32
// master): per layer W[Co][taps][Ci] then b[Co], layers in the order above.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "guess.h"

using namespace chfem::nn;

namespace {

struct Layer {
    int mode, Ci, Co, level_in, level_out, relu, skip;  // skip: activation index added before relu (-1 none)
    size_t woff, boff, nw;                               // offsets in the parameter vector
    uint16_t *wf = nullptr, *wb = nullptr;               // packed forward / input-gradient weights
};

// activation slots
enum { A_IN, A_E0A, A_E0, A_E1A, A_E1, A_E2A, A_E2, A_E3A, A_E3, A_D2, A_D1, A_D0, A_COUNT };

struct Net {
    int physics, nout, Cout;
    std::vector<Layer> layers;
    std::vector<int> act_in, act_out;  // per layer: input and output activation slot (output of last: y)
    int act_level[A_COUNT], act_ch[A_COUNT];
    size_t nparams = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // device FP32
    int step = 0;
    hipStream_t s = nullptr;
    // per-grid buffers
    Grid3 grid[4] = {};
    uint16_t *act[A_COUNT] = {};
    float *y = nullptr;
    float *gact[A_COUNT] = {};
    float *gy = nullptr, *tmp = nullptr, *wpart = nullptr, *bpart = nullptr;
    uint16_t *dz = nullptr;
    double *lpart = nullptr, *loss = nullptr;
    float *target = nullptr;
    uint16_t *elem = nullptr;
    float *table = nullptr;
    size_t cap_voxels = 0, cap_train = 0, cap_target = 0;
    bool packed = false;
};

bool ok(hipError_t e)
{
    if (e != hipSuccess) {
        fprintf(stderr, "chfem guess: %s\n", hipGetErrorString(e));
        return false;
    }
    return true;
}

template <typename T> void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void add_layer(Net &n, int mode, int Ci, int Co, int in, int out, int relu, int skip)
{
    Layer L;
    L.mode = mode;
    L.Ci = Ci;
    L.Co = Co;
    L.relu = relu;
    L.skip = skip;
    const int taps = mode == CONV_DOWN ? 8 : 27;
    L.nw = (size_t)Co * taps * Ci;
    L.woff = n.nparams;
    L.boff = L.woff + L.nw;
    n.nparams = L.boff + Co;
    L.level_in = in >= 0 ? n.act_level[in] : 0;
    L.level_out = out >= 0 ? n.act_level[out] : 0;
    n.layers.push_back(L);
    n.act_in.push_back(in);
    n.act_out.push_back(out);
}

void build(Net &n)
{
    const int lv[A_COUNT] = {0, 0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0};
    const int ch[A_COUNT] = {16, 16, 16, 32, 32, 48, 48, 48, 48, 48, 32, 16};
    memcpy(n.act_level, lv, sizeof(lv));
    memcpy(n.act_ch, ch, sizeof(ch));
    add_layer(n, CONV_SAME, 16, 16, A_IN, A_E0A, 1, -1);
    add_layer(n, CONV_SAME, 16, 16, A_E0A, A_E0, 1, -1);
    add_layer(n, CONV_DOWN, 16, 32, A_E0, A_E1A, 1, -1);
    add_layer(n, CONV_SAME, 32, 32, A_E1A, A_E1, 1, -1);
    add_layer(n, CONV_DOWN, 32, 48, A_E1, A_E2A, 1, -1);
    add_layer(n, CONV_SAME, 48, 48, A_E2A, A_E2, 1, -1);
    add_layer(n, CONV_DOWN, 48, 48, A_E2, A_E3A, 1, -1);
    add_layer(n, CONV_SAME, 48, 48, A_E3A, A_E3, 1, -1);
    add_layer(n, CONV_UP, 48, 48, A_E3, A_D2, 1, A_E2);
    add_layer(n, CONV_UP, 48, 32, A_D2, A_D1, 1, A_E1);
    add_layer(n, CONV_UP, 32, 16, A_D1, A_D0, 1, A_E0);
    add_layer(n, CONV_SAME, 16, n.Cout, A_D0, -1, 0, -1);
}

// splitmix64 -> uniform in [0, 1)
double urand(uint64_t &st)
{
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

bool ensure_grids(Net &n, int nx, int ny, int nz, bool training)
{
    const size_t V = (size_t)nx * ny * nz;
    for (int k = 0; k < 4; k++) n.grid[k] = Grid3{ny >> k, nx >> k, nz >> k};
    if (V > n.cap_voxels) {
        for (int a = 0; a < A_COUNT; a++) dfree(n.act[a]);
        dfree(n.y);
        dfree(n.elem);
        n.cap_voxels = 0;
        for (int a = 0; a < A_COUNT; a++) {
            const size_t na = (V >> (3 * n.act_level[a])) * n.act_ch[a];
            if (!ok(hipMalloc(&n.act[a], na * sizeof(uint16_t)))) return false;
        }
        if (!ok(hipMalloc(&n.y, V * n.Cout * sizeof(float)))) return false;
        if (!ok(hipMalloc(&n.elem, V * sizeof(uint16_t)))) return false;
        n.cap_voxels = V;
    }
    if (!training) return true;
    if (V > n.cap_train) {
        for (int a = 0; a < A_COUNT; a++) dfree(n.gact[a]);
        dfree(n.gy); dfree(n.tmp); dfree(n.wpart); dfree(n.bpart); dfree(n.dz); dfree(n.lpart);
        n.cap_train = 0;
        for (int a = 1; a < A_COUNT; a++) {
            const size_t na = (V >> (3 * n.act_level[a])) * n.act_ch[a];
            if (!ok(hipMalloc(&n.gact[a], na * sizeof(float)))) return false;
        }
        if (!ok(hipMalloc(&n.gy, V * n.Cout * sizeof(float)))) return false;
        // input gradient of an up-convolution before pooling: fine grid, Ci channels
        size_t tmpn = 0, wp = 0, bp = 0, dzn = 0;
        for (const Layer &L : n.layers) {
            const Grid3 go = n.grid[L.level_out];
            const size_t Vo = (size_t)go.ny * go.nx * go.nz;
            if (L.mode == CONV_UP) tmpn = tmpn > Vo * L.Ci ? tmpn : Vo * L.Ci;
            const size_t w = wpart_size(go, L.Co, L.Ci, L.mode == CONV_DOWN ? 8 : 27);
            wp = wp > w ? wp : w;
            const size_t b = bias_part_size(Vo, L.Co);
            bp = bp > b ? bp : b;
            dzn = dzn > Vo * L.Co ? dzn : Vo * L.Co;
        }
        if (!ok(hipMalloc(&n.tmp, tmpn * sizeof(float)))) return false;
        if (!ok(hipMalloc(&n.wpart, wp * sizeof(float)))) return false;
        if (!ok(hipMalloc(&n.bpart, bp * sizeof(float)))) return false;
        if (!ok(hipMalloc(&n.dz, dzn * sizeof(uint16_t)))) return false;
        if (!ok(hipMalloc(&n.lpart, (loss_part_size(V) + 1) * sizeof(double)))) return false;
        n.cap_train = V;
    }
    return true;
}

bool pack(Net &n)
{
    if (n.packed) return true;
    for (Layer &L : n.layers) {
        const int taps = L.mode == CONV_DOWN ? 8 : 27;
        if (!L.wf && !ok(hipMalloc(&L.wf, L.nw * sizeof(uint16_t)))) return false;
        if (!L.wb && !ok(hipMalloc(&L.wb, L.nw * sizeof(uint16_t)))) return false;
        pack_weights(n.w + L.woff, L.wf, L.Co, L.Ci, taps, 0, n.s);
        pack_weights(n.w + L.woff, L.wb, L.Co, L.Ci, taps, L.mode == CONV_DOWN ? 2 : 1, n.s);
    }
    n.packed = true;
    return true;
}

// material feature table: log10(p0 / max p0) and (nu - 0.25) * 4 (elastic)
bool set_input(Net &n, const uint16_t *elem, const double *props, unsigned nmat, size_t V)
{
    for (size_t v = 0; v < V; v++)
        if (elem[v] >= nmat) {
            fprintf(stderr, "chfem guess: element %zu has material %u of %u\n", v, (unsigned)elem[v], nmat);
            return false;
        }
    std::vector<float> tab(2 * (size_t)nmat, 0.f);
    double pmax = 0;
    for (unsigned k = 0; k < nmat; k++) {
        const double p = n.physics == 1 ? props[2 * k] : props[k];
        pmax = p > pmax ? p : pmax;
    }
    if (!(pmax > 0)) return false;
    for (unsigned k = 0; k < nmat; k++) {
        const double p = n.physics == 1 ? props[2 * k] : props[k];
        tab[2 * k] = p > 0 ? (float)log10(p / pmax) : -6.f;
        tab[2 * k + 1] = n.physics == 1 ? (float)((props[2 * k + 1] - 0.25) * 4.0) : 0.f;
    }
    dfree(n.table);
    if (!ok(hipMalloc(&n.table, tab.size() * sizeof(float)))) return false;
    if (!ok(hipMemcpyAsync(n.table, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice, n.s))) return false;
    if (!ok(hipMemcpyAsync(n.elem, elem, V * sizeof(uint16_t), hipMemcpyHostToDevice, n.s))) return false;
    features(n.elem, n.table, 2, n.act[A_IN], 16, V, n.s);
    return ok(hipStreamSynchronize(n.s));
}

void forward(Net &n)
{
    pack(n);
    for (size_t k = 0; k < n.layers.size(); k++) {
        const Layer &L = n.layers[k];
        const Grid3 gi = n.grid[L.level_in], go = n.grid[L.level_out];
        const uint16_t *skip = L.skip >= 0 ? n.act[L.skip] : nullptr;
        const bool last = n.act_out[k] < 0;
        conv(L.mode, n.act[n.act_in[k]], gi, L.Ci, L.wf, n.w + L.boff, skip, L.relu,
             last ? (void *)n.y : (void *)n.act[n.act_out[k]], last, 0, go, L.Co, n.s);
    }
}

// gradients of the loss (dy in n.gy) accumulated into n.g
void backward(Net &n)
{
    // skip targets receive their first contribution through act_backward: zero them
    const int zero_first[3] = {A_E0, A_E1, A_E2};
    for (int a : zero_first) {
        const Grid3 gg = n.grid[n.act_level[a]];
        fill_f32(n.gact[a], 0.f, (size_t)gg.ny * gg.nx * gg.nz * n.act_ch[a], n.s);
    }
    for (int k = (int)n.layers.size() - 1; k >= 0; k--) {
        const Layer &L = n.layers[k];
        const Grid3 gi = n.grid[L.level_in], go = n.grid[L.level_out];
        const size_t Vo = (size_t)go.ny * go.nx * go.nz;
        const int out = n.act_out[k], in = n.act_in[k];
        const float *dy = out < 0 ? n.gy : n.gact[out];
        const uint16_t *yv = out < 0 ? nullptr : n.act[out];
        act_backward(dy, yv, L.relu, n.dz, L.skip >= 0 ? n.gact[L.skip] : nullptr, n.bpart, n.g + L.boff, Vo, L.Co,
                     n.s);
        conv_wgrad(L.mode, n.act[in], gi, L.Ci, n.dz, go, L.Co, n.wpart, n.g + L.woff, n.s);
        if (in == A_IN) continue;
        float *gin = n.gact[in];
        // inputs that are skip sources (e0, e1, e2) already hold their skip gradient
        const bool acc = in == A_E0 || in == A_E1 || in == A_E2;
        if (L.mode == CONV_SAME) {
            conv(CONV_SAME, n.dz, go, L.Co, L.wb, nullptr, nullptr, 0, gin, 1, acc, gi, L.Ci, n.s);
        } else if (L.mode == CONV_DOWN) {
            conv(CONV_PARITY, n.dz, go, L.Co, L.wb, nullptr, nullptr, 0, gin, 1, acc, gi, L.Ci, n.s);
        } else {
            conv(CONV_SAME, n.dz, go, L.Co, L.wb, nullptr, nullptr, 0, n.tmp, 1, 0, go, L.Ci, n.s);
            pool_children(n.tmp, go, gin, L.Ci, acc, n.s);
        }
    }
}

Net *as_net(void *h) { return (Net *)h; }

bool dims_ok(unsigned nx, unsigned ny, unsigned nz) { return nx % 8 == 0 && ny % 8 == 0 && nz % 8 == 0 && nx && ny && nz; }

}  // namespace

namespace chfem {
// Used by femhmg.cpp: prediction of the fluctuation fields of a model.
void *guess_active = nullptr;

// Runs the network on an element map (host) and leaves the prediction on the
// device; returns the device pointer to y[V][Cout] or nullptr.
const float *guess_predict_device(void *h, const uint16_t *elem_host, const double *props, unsigned nmat, unsigned nx,
                                  unsigned ny, unsigned nz, int *Cout)
{
    Net *n = as_net(h);
    if (!n || !dims_ok(nx, ny, nz)) return nullptr;
    if (!ensure_grids(*n, (int)nx, (int)ny, (int)nz, false)) return nullptr;
    if (!set_input(*n, elem_host, props, nmat, (size_t)nx * ny * nz)) return nullptr;
    forward(*n);
    *Cout = n->Cout;
    return n->y;
}

hipStream_t guess_stream(void *h) { return h ? as_net(h)->s : nullptr; }
int guess_physics(void *h) { return h ? as_net(h)->physics : -1; }
}  // namespace chfem

extern "C" {

void chfemGuessDestroy(void *h);

void *chfemGuessCreate(int physics, unsigned seed)
{
    if (physics != 0 && physics != 1) return nullptr;
    Net *n = new Net();
    n->physics = physics;
    n->nout = physics == 1 ? 18 : 3;
    n->Cout = physics == 1 ? 32 : 16;
    build(*n);
    if (!ok(hipStreamCreate(&n->s))) {
        delete n;
        return nullptr;
    }
    // He-uniform init of the real input channels; padded channels stay zero
    std::vector<float> w(n->nparams, 0.f);
    uint64_t st = 0x5eedull + seed;
    for (size_t k = 0; k < n->layers.size(); k++) {
        const Layer &L = n->layers[k];
        const int taps = L.mode == CONV_DOWN ? 8 : 27;
        const int ci_real = k == 0 ? 2 : L.Ci;
        const int co_real = k + 1 == n->layers.size() ? n->nout : L.Co;
        const double bound = sqrt(6.0 / (taps * ci_real)) * (k + 1 == n->layers.size() ? 0.1 : 1.0);
        for (int co = 0; co < co_real; co++)
            for (int t = 0; t < taps; t++)
                for (int ci = 0; ci < ci_real; ci++)
                    w[L.woff + ((size_t)co * taps + t) * L.Ci + ci] = (float)((2.0 * urand(st) - 1.0) * bound);
    }
    const size_t b = n->nparams * sizeof(float);
    if (!ok(hipMalloc(&n->w, b)) || !ok(hipMalloc(&n->g, b)) || !ok(hipMalloc(&n->m, b)) || !ok(hipMalloc(&n->v, b)) ||
        !ok(hipMalloc(&n->loss, sizeof(double)))) {
        chfemGuessDestroy(n);
        return nullptr;
    }
    if (!ok(hipMemcpy(n->w, w.data(), b, hipMemcpyHostToDevice)) || !ok(hipMemset(n->g, 0, b)) ||
        !ok(hipMemset(n->m, 0, b)) || !ok(hipMemset(n->v, 0, b))) {
        chfemGuessDestroy(n);
        return nullptr;
    }
    return n;
}

void chfemGuessDestroy(void *h)
{
    Net *n = as_net(h);
    if (!n) return;
    if (chfem::guess_active == h) chfem::guess_active = nullptr;
    (void)hipStreamSynchronize(n->s);
    for (Layer &L : n->layers) {
        dfree(L.wf);
        dfree(L.wb);
    }
    for (int a = 0; a < A_COUNT; a++) {
        dfree(n->act[a]);
        dfree(n->gact[a]);
    }
    dfree(n->y); dfree(n->gy); dfree(n->tmp); dfree(n->wpart); dfree(n->bpart); dfree(n->dz); dfree(n->lpart);
    dfree(n->loss); dfree(n->target); dfree(n->elem); dfree(n->table);
    dfree(n->w); dfree(n->g); dfree(n->m); dfree(n->v);
    (void)hipStreamDestroy(n->s);
    delete n;
}

size_t chfemGuessNumParams(void *h) { return h ? as_net(h)->nparams : 0; }
unsigned chfemGuessNumOutputs(void *h) { return h ? (unsigned)as_net(h)->nout : 0; }

int chfemGuessGetParams(void *h, float *out)
{
    Net *n = as_net(h);
    if (!n || !out) return 0;
    return ok(hipMemcpy(out, n->w, n->nparams * sizeof(float), hipMemcpyDeviceToHost));
}

int chfemGuessSetParams(void *h, const float *in)
{
    Net *n = as_net(h);
    if (!n || !in) return 0;
    n->packed = false;
    return ok(hipMemcpy(n->w, in, n->nparams * sizeof(float), hipMemcpyHostToDevice));
}

int chfemGuessGetGradients(void *h, float *out)
{
    Net *n = as_net(h);
    if (!n || !out) return 0;
    if (!ok(hipStreamSynchronize(n->s))) return 0;
    return ok(hipMemcpy(out, n->g, n->nparams * sizeof(float), hipMemcpyDeviceToHost));
}

int chfemGuessZeroGradients(void *h)
{
    Net *n = as_net(h);
    if (!n) return 0;
    fill_f32(n->g, 0.f, n->nparams, n->s);
    return ok(hipStreamSynchronize(n->s));
}

int chfemGuessPredict(void *h, const uint16_t *elem, const double *props, unsigned nmat, unsigned nx, unsigned ny,
                      unsigned nz, float *out)
{
    Net *n = as_net(h);
    if (!n || !elem || !props || !out || !nmat || nmat > 256) return 0;
    int C = 0;
    const float *y = chfem::guess_predict_device(h, elem, props, nmat, nx, ny, nz, &C);
    if (!y) return 0;
    const size_t V = (size_t)nx * ny * nz;
    std::vector<float> full(V * C);
    if (!ok(hipMemcpyAsync(full.data(), y, full.size() * sizeof(float), hipMemcpyDeviceToHost, n->s))) return 0;
    if (!ok(hipStreamSynchronize(n->s))) return 0;
    for (size_t v = 0; v < V; v++) memcpy(out + v * n->nout, full.data() + v * C, n->nout * sizeof(float));
    return 1;
}

int chfemGuessAccumulate(void *h, const uint16_t *elem, const double *props, unsigned nmat, unsigned nx, unsigned ny,
                         unsigned nz, const float *target, double *loss)
{
    Net *n = as_net(h);
    if (!n || !elem || !props || !target || !nmat || nmat > 256 || !dims_ok(nx, ny, nz)) return 0;
    const size_t V = (size_t)nx * ny * nz;
    if (!ensure_grids(*n, (int)nx, (int)ny, (int)nz, true)) return 0;
    if (!set_input(*n, elem, props, nmat, V)) return 0;
    const size_t nt = V * n->nout;
    if (nt > n->cap_target) {
        dfree(n->target);
        if (!ok(hipMalloc(&n->target, nt * sizeof(float)))) return 0;
        n->cap_target = nt;
    }
    if (!ok(hipMemcpyAsync(n->target, target, nt * sizeof(float), hipMemcpyHostToDevice, n->s))) return 0;
    forward(*n);
    if (!ok(hipMemsetAsync(n->loss, 0, sizeof(double), n->s))) return 0;
    mse(n->y, n->Cout, n->target, n->nout, V, 1.0 / (double)nt, n->gy, n->lpart, n->loss, n->s);
    backward(*n);
    double l = 0;
    if (!ok(hipMemcpyAsync(&l, n->loss, sizeof(double), hipMemcpyDeviceToHost, n->s))) return 0;
    if (!ok(hipStreamSynchronize(n->s))) return 0;
    if (loss) *loss = l;
    return 1;
}

int chfemGuessStep(void *h, float lr, float grad_scale)
{
    Net *n = as_net(h);
    if (!n) return 0;
    n->step++;
    adam(n->w, n->g, n->m, n->v, n->nparams, lr, 0.9f, 0.999f, 1e-8f, n->step, grad_scale, n->s);
    fill_f32(n->g, 0.f, n->nparams, n->s);
    n->packed = false;
    return ok(hipStreamSynchronize(n->s));
}

int chfemGuessMfmaCheck(const uint16_t *A, const uint16_t *B, float *D)
{
    uint16_t *dA = nullptr, *dB = nullptr;
    float *dD = nullptr;
    int r = ok(hipMalloc(&dA, 512)) && ok(hipMalloc(&dB, 512)) && ok(hipMalloc(&dD, 1024)) &&
            ok(hipMemcpy(dA, A, 512, hipMemcpyHostToDevice)) && ok(hipMemcpy(dB, B, 512, hipMemcpyHostToDevice));
    if (r) {
        mfma_check(dA, dB, dD, nullptr);
        r = ok(hipDeviceSynchronize()) && ok(hipMemcpy(D, dD, 1024, hipMemcpyDeviceToHost));
    }
    dfree(dA);
    dfree(dB);
    dfree(dD);
    return r;
}

int hmgSetInitialGuessNetwork(void *h)
{
    if (h && as_net(h)->physics != 0 && as_net(h)->physics != 1) return 0;
    chfem::guess_active = h;
    return 1;
}

}  // extern "C"
