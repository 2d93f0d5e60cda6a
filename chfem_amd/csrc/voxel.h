// voxel.h -- local matrices of the unit-voxel trilinear hexahedron from exact
// 1D integrals of the linear shape functions (mass 1/3 and 1/6, stiffness +-1,
// mixed +-1/2).  Local node k of an element sits at (row + DR[k], col + DC[k],
// layer + DL[k]) (chfem/src/femhmg/physical_phenomena/femhmg_3D.c:26-73); the
// element frame is x = col offset, y = 1 - row offset and z = 1 - layer offset
// for thermal/elastic, z = layer offset for Stokes.
#pragma once

namespace chfem {
namespace voxel {

static const int DR[8] = {1, 1, 0, 0, 1, 1, 0, 0};
static const int DC[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int DL[8] = {0, 0, 0, 0, 1, 1, 1, 1};

inline void frame(int k, bool stokes, int &x, int &y, int &z)
{
    x = DC[k];
    y = 1 - DR[k];
    z = stokes ? DL[k] : 1 - DL[k];
}

// int over the unit cube of d_p N_i * d_q N_j  (p or q = -1: no derivative)
inline double int_dd(int i, int j, int p, int q, bool stokes)
{
    int ci[3], cj[3];
    frame(i, stokes, ci[0], ci[1], ci[2]);
    frame(j, stokes, cj[0], cj[1], cj[2]);
    double v = 1.0;
    for (int t = 0; t < 3; t++) {
        const double si = ci[t] ? 1.0 : -1.0, sj = cj[t] ? 1.0 : -1.0;
        if (t == p && t == q) v *= si * sj;
        else if (t == p) v *= 0.5 * si;
        else if (t == q) v *= 0.5 * sj;
        else v *= ci[t] == cj[t] ? 1.0 / 3.0 : 1.0 / 6.0;
    }
    return v;
}

inline double lap(int i, int j, bool stokes)
{
    return int_dd(i, j, 0, 0, stokes) + int_dd(i, j, 1, 1, stokes) + int_dd(i, j, 2, 2, stokes);
}

// 24x24 isotropic stiffness lambda*A + mu*B, DOFs node-major (x,y,z)
inline void elastic_K(double lam, double mu, bool stokes, double *K)
{
    for (int i = 0; i < 8; i++)
        for (int a = 0; a < 3; a++)
            for (int j = 0; j < 8; j++)
                for (int b = 0; b < 3; b++)
                    K[(3 * i + a) * 24 + 3 * j + b] =
                        lam * int_dd(i, j, a, b, stokes) +
                        mu * ((a == b ? lap(i, j, stokes) : 0.0) + int_dd(i, j, b, a, stokes));
}

// (lambda, mu) of a 24x24 matrix if it is an isotropic unit-voxel stiffness:
// K[0x,0x] = (lambda + 4 mu)/9 and K[0x,0y] = (lambda + mu)/12 for node 0.
inline bool isotropic_params(const double *K, double &lam, double &mu)
{
    const double a = 9.0 * K[0], b = 12.0 * K[1];
    mu = (a - b) / 3.0;
    lam = b - mu;
    double R[576], kmax = 0.0, dmax = 0.0;
    elastic_K(lam, mu, false, R);
    for (int i = 0; i < 576; i++) {
        kmax = fabs(K[i]) > kmax ? fabs(K[i]) : kmax;
        const double d = fabs(K[i] - R[i]);
        dmax = d > dmax ? d : dmax;
    }
    return kmax > 0.0 && dmax <= 1e-12 * kmax;
}

// Conductivity k of an 8x8 matrix if it is the unit-voxel trilinear conduction
// matrix k * {1/3 on the diagonal, 0 between nodes that differ in one
// coordinate, -1/12 between nodes that differ in two or three}
// (local node order of DR/DC/DL).
inline bool conduction_param(const double *K, double &k)
{
    static const int dr[8] = {1, 1, 0, 0, 1, 1, 0, 0}, dc[8] = {0, 1, 1, 0, 0, 1, 1, 0}, dl[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    k = 3.0 * K[0];
    double kmax = 0.0, dmax = 0.0;
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 8; j++) {
            const int diff = (dr[i] != dr[j]) + (dc[i] != dc[j]) + (dl[i] != dl[j]);
            const double r = diff == 0 ? k / 3.0 : (diff == 1 ? 0.0 : -k / 12.0);
            const double v = K[i * 8 + j];
            kmax = fabs(v) > kmax ? fabs(v) : kmax;
            dmax = fabs(v - r) > dmax ? fabs(v - r) : dmax;
        }
    return kmax > 0.0 && dmax <= 1e-12 * kmax;
}

// Inverse Jacobi diagonal of every 16-bit packed key (2 bits per local node,
// at most 4 materials): 1 / sum_e K[m_e][e][a] summed over e = 0..7 in FP64,
// at [key * nd + a]; the diagonal itself (the sum) at [65536 * nd + key * nd + a].
// K: dense (8 nd)^2 per material; entries of keys with a material >= nmat
// stay 0.  t64/t32 get 2 * 65536 * nd entries.
template <typename V64, typename V32>
inline void inverse_diagonal_table(const double *K, unsigned nmat, int nd, V64 &t64, V32 &t32)
{
    const int ke = 8 * nd;
    const size_t ksz = (size_t)ke * ke;
    const size_t half = (size_t)65536 * nd;
    t64.assign(2 * half, 0.0);
    t32.assign(2 * half, 0.f);
    for (unsigned key = 0; key < 65536; key++) {
        bool valid = true;
        for (int e = 0; e < 8; e++) valid = valid && ((key >> (2 * e)) & 3u) < nmat;
        if (!valid) continue;
        for (int a = 0; a < nd; a++) {
            double sum = 0;
            for (int e = 0; e < 8; e++) sum += K[(size_t)((key >> (2 * e)) & 3u) * ksz + (size_t)(nd * e + a) * ke + nd * e + a];
            t64[(size_t)key * nd + a] = 1.0 / sum;
            t32[(size_t)key * nd + a] = (float)(1.0 / sum);
            t64[half + (size_t)key * nd + a] = sum;
            t32[half + (size_t)key * nd + a] = (float)sum;
        }
    }
}

}  // namespace voxel
}  // namespace chfem
