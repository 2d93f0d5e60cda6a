/*
 * chfem_oracle.c -- CPU FP64 restatement of cortezpedro/chfem's matrix-free
 * homogenization path for 3D voxel images: thermal (1 DOF/node), linear
 * elastic (3 DOF/node) and stabilised Stokes (4 DOF/node) with periodic
 * boundaries.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the parity checker for the HIP product
 * library and the timed CPU baseline.  Only the test-suite, the smoke check and
 * the benchmark's CPU-baseline leg load it; the product never links or calls
 * it.  It follows the reference's algorithm step by step, in plain loops, so
 * that a reader can hold each function next to the reference lines it cites
 * (paths relative to the reference repository root).
 *
 * Conventions restated from the reference:
 *   - rows = ny (row 0 at the TOP), cols = nx, layers = nz, counted in
 *     elements.  Element e and periodic node i share the index
 *     i = row + col*ny + layer*nx*ny  (chfem/src/femhmg/cudapcg/kernels/utils.h:81-82,
 *     chfem/src/femhmg/physical_phenomena/femhmg_3D.c:4-17).
 *   - local node k of element (r,c,l) sits at (r+DR[k], c+DC[k], l+DL[k]):
 *     0 left-bottom-near, 1 right-bottom-near, 2 right-top-near,
 *     3 left-top-near, 4..7 the same on the far face
 *     (femhmg_3D.c:26-73).
 *   - element frames: x = col offset, y = 1-row offset (y points up); for
 *     thermal/elastic z = 1-layer offset (z points to the near face,
 *     femhmg_elastic_3D.c:51-128 sign pattern, femhmg_thermal_3D.c:438-492),
 *     for Stokes z = layer offset (femhmg_fluid_3D.c:41-110 sign pattern).
 *
 * Parity pins (see tests/): the three 3D golden results of the reference's
 * own test-suite, homogeneous-image closed forms, laminate Voigt/Reuss
 * bounds, C symmetry, axis-permutation covariance, brute-force assembled
 * matrices on tiny grids, rigid-body null spaces and the Stokes stencil
 * constants.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define ORC_THERMAL 0
#define ORC_ELASTIC 1
#define ORC_FLUID   2

static const int DR[8] = {1, 1, 0, 0, 1, 1, 0, 0};
static const int DC[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int DL[8] = {0, 0, 0, 0, 1, 1, 1, 1};

/* load case index of the thermal expansion analysis (HOMOGENIZE_THERMAL_EXPANSION
 * is 7 in chfem/src/femhmg/includes.h:45; here it follows the six seam cases) */
#define ORC_DIR_TEXP 6

/* periodic node/element index, PERIODICNUM_3D (utils.h:82) */
static inline uint32_t pidx(long r, long c, long l, long ny, long nx, long nz)
{
    r %= ny; if (r < 0) r += ny;
    c %= nx; if (c < 0) c += nx;
    l %= nz; if (l < 0) l += nz;
    return (uint32_t)(r + c * ny + l * nx * ny);
}


/* The 27 periodic node indices around node/element i, nb[a + 3b + 9c] for
 * row, column and layer offsets a-1, b-1, c-1 (PERIODICNUM_3D, utils.h:82,
 * with the wrap done by comparison instead of a modulo per neighbour). */
static inline void nbr27(size_t i, long ny, long nx, long nz, uint32_t nb[27])
{
    long r = (long)(i % (size_t)ny), c = (long)((i / (size_t)ny) % (size_t)nx), l = (long)(i / ((size_t)nx * ny));
    long R[3] = {r == 0 ? ny - 1 : r - 1, r, r + 1 == ny ? 0 : r + 1};
    long Cc[3] = {(c == 0 ? nx - 1 : c - 1) * ny, c * ny, (c + 1 == nx ? 0 : c + 1) * ny};
    long Ll[3] = {(l == 0 ? nz - 1 : l - 1) * nx * ny, l * nx * ny, (l + 1 == nz ? 0 : l + 1) * nx * ny};
    for (int cl = 0; cl < 3; cl++)
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) nb[a + 3 * b + 9 * cl] = (uint32_t)(R[a] + Cc[b] + Ll[cl]);
}
/* offset slot of local node j of the element for which the centre node is
 * local node e (element origin = node - (DR,DC,DL)[e]) */
#define NB_SLOT(e, j) ((1 + DR[j] - DR[e]) + 3 * (1 + DC[j] - DC[e]) + 9 * (1 + DL[j] - DL[e]))
/* slot of local node j of the element whose origin is the centre */
#define NB_ELEM(j) ((1 + DR[j]) + 3 * (1 + DC[j]) + 9 * (1 + DL[j]))
/* slot of the origin of the element for which the centre is local node e */
#define NB_ORIG(e) ((1 - DR[e]) + 3 * (1 - DC[e]) + 9 * (1 - DL[e]))

/* =========================================================================
 * A1: image -> element material map with colour keys, x<->y transpose and
 *     mesh refinement.  Restates readMaterialMapNumpy / readMaterialMapRAW
 *     (chfem/src/femhmg/femhmg.c:1711-1758, :1644-1709).
 *     img is in file order (z slices, y rows top->bottom, x fastest) with
 *     dimensions (nx0, ny0, nz0) before refinement.
 * ========================================================================= */
void orc_image_to_elem_map(const uint8_t *img, unsigned nx0, unsigned ny0, unsigned nz0,
                           unsigned ref, const uint8_t *props_keys, uint16_t *elem)
{
    const size_t rows_ref = (size_t)ny0 * ref, cols_ref = (size_t)nx0 * ref;
    size_t idx = 0;
    for (unsigned k = 0; k < nz0; k++)
        for (unsigned i = 0; i < ny0; i++)
            for (unsigned j = 0; j < nx0; j++) {
                uint16_t m = props_keys[img[idx++]];
                for (unsigned kk = ref * k; kk < ref * (k + 1); kk++)
                    for (unsigned ii = ref * i; ii < ref * (i + 1); ii++)
                        for (unsigned jj = ref * j; jj < ref * (j + 1); jj++)
                            elem[ii + jj * rows_ref + kk * rows_ref * cols_ref] = m;
            }
}

/* A2: node on the (ny+1)x(nx+1)x(nz+1) grid -> periodic DOF
 *     (femhmg_3D.c:4-17). */
void orc_node_dof_map(unsigned nx, unsigned ny, unsigned nz, unsigned nnodedof, uint32_t *map)
{
    const size_t NY = ny + 1, NX = nx + 1, NZ = nz + 1;
    for (size_t l = 0; l < NZ; l++)
        for (size_t c = 0; c < NX; c++)
            for (size_t r = 0; r < NY; r++)
                map[r + c * NY + l * NX * NY] = pidx(r, c, l, ny, nx, nz) * nnodedof;
}

/* A3: packed 16-bit node key, bits [2k,2k+1] = material of the element for
 *     which the node is local node k (femhmg_3D.c:19-75, MATKEY_BITSTEP_3D=2,
 *     chfem/src/femhmg/cudapcg/includes.h:77-83). */
void orc_dof_material_map(unsigned nx, unsigned ny, unsigned nz, const uint16_t *elem, uint16_t *keys)
{
    const size_t N = (size_t)nx * ny * nz;
    memset(keys, 0, N * sizeof(uint16_t));
    for (int k = 0; k < 8; k++)
        for (size_t e = 0; e < N; e++) {
            long r = e % ny, c = (e / ny) % nx, l = e / ((size_t)nx * ny);
            keys[pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz)] += (uint16_t)((elem[e] & 3u) << (2 * k));
        }
}

/* A4: fluid node keys, bit k set iff the element with the node as local node
 *     k is pore (material index 0) (femhmg_fluid_3D.c:426-491), and the
 *     DOF-number pore mapping (femhmg_fluid_3D.c:370-424).  Returns the
 *     number of pore elements. */
size_t orc_fluid_keys(unsigned nx, unsigned ny, unsigned nz, const uint16_t *elem, uint8_t *fkeys)
{
    const size_t N = (size_t)nx * ny * nz;
    size_t npore = 0;
    for (size_t e = 0; e < N; e++) npore += (elem[e] == 0);
    memset(fkeys, 0, N);
    for (int k = 0; k < 8; k++)
        for (size_t e = 0; e < N; e++) {
            long r = e % ny, c = (e / ny) % nx, l = e / ((size_t)nx * ny);
            fkeys[pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz)] += (uint8_t)((elem[e] < 1) << k);
        }
    return npore;
}

void orc_fluid_dof_ids(size_t N, const uint8_t *fkeys, uint32_t *dof_id, uint32_t *nvel, uint32_t *nborder,
                       uint32_t *dof2node /* nullable, size nvel+nborder */, uint8_t *border_keys /* nullable */)
{
    uint32_t count = 0;
    for (size_t i = 0; i < N; i++) {
        if (fkeys[i] == 255) dof_id[i] = count++;
        else dof_id[i] = 0;
    }
    *nvel = count;
    for (size_t i = 0; i < N; i++)
        if (fkeys[i] != 255 && fkeys[i] != 0) dof_id[i] = count++;
    *nborder = count - *nvel;
    if (dof2node)
        for (size_t i = 0; i < N; i++)
            if (fkeys[i]) dof2node[dof_id[i]] = (uint32_t)i;
    if (border_keys)
        for (size_t i = 0; i < N; i++)
            if (fkeys[i] && fkeys[i] != 255) border_keys[dof_id[i] - *nvel] = fkeys[i];
}

/* =========================================================================
 * A5: local matrices of a unit voxel by exact 2x2x2 Gauss quadrature of the
 *     trilinear hexahedron.  The reference stores closed forms
 *     (femhmg_thermal_3D.c:42-120, femhmg_elastic_3D.c:51-128,
 *     femhmg_fluid_3D.c:41-110); the integrals below are the same bilinear
 *     forms evaluated from the shape functions.
 * ========================================================================= */
static void node_xyz(int k, int fluid_frame, double *x, double *y, double *z)
{
    *x = DC[k];
    *y = 1 - DR[k];
    *z = fluid_frame ? DL[k] : 1 - DL[k];
}

/* grad N_k at (p,q,s) in the unit cube */
static void grad_shape(int k, int fluid_frame, double p, double q, double s, double g[3], double *N)
{
    double xk, yk, zk;
    node_xyz(k, fluid_frame, &xk, &yk, &zk);
    double fx = xk > 0.5 ? p : 1 - p, dfx = xk > 0.5 ? 1 : -1;
    double fy = yk > 0.5 ? q : 1 - q, dfy = yk > 0.5 ? 1 : -1;
    double fz = zk > 0.5 ? s : 1 - s, dfz = zk > 0.5 ? 1 : -1;
    g[0] = dfx * fy * fz;
    g[1] = fx * dfy * fz;
    g[2] = fx * fy * dfz;
    if (N) *N = fx * fy * fz;
}

static const double GP[2] = {0.21132486540518711775, 0.78867513459481288225};

/* dense 8x8 conduction matrix for unit conductivity */
void orc_ke_thermal_unit(double *K64)
{
    memset(K64, 0, 64 * sizeof(double));
    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
        double g[8][3];
        for (int k = 0; k < 8; k++) grad_shape(k, 0, GP[a], GP[b], GP[c], g[k], NULL);
        for (int i = 0; i < 8; i++) for (int j = 0; j < 8; j++)
            K64[i * 8 + j] += 0.125 * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
    }
}

/* Voigt strain-displacement matrix B (6x24), rows xx,yy,zz,yz,xz,xy with
 * engineering shears, columns node-major (x,y,z). */
static void bmat(int fluid_frame, double p, double q, double s, double B[6][24])
{
    memset(B, 0, sizeof(double) * 6 * 24);
    for (int k = 0; k < 8; k++) {
        double g[3];
        grad_shape(k, fluid_frame, p, q, s, g, NULL);
        B[0][3 * k + 0] = g[0];
        B[1][3 * k + 1] = g[1];
        B[2][3 * k + 2] = g[2];
        B[3][3 * k + 1] = g[2]; B[3][3 * k + 2] = g[1];
        B[4][3 * k + 0] = g[2]; B[4][3 * k + 2] = g[0];
        B[5][3 * k + 0] = g[1]; B[5][3 * k + 1] = g[0];
    }
}

static void dmat(double lam, double mu, double D[6][6])
{
    memset(D, 0, sizeof(double) * 36);
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) D[i][j] = lam;
        D[i][i] = lam + 2 * mu;
        D[3 + i][3 + i] = mu;
    }
}

/* 24x24 stiffness and 6x24 centre stress operator CB = D B(1/2,1/2,1/2) */
void orc_ke_elastic(double E, double nu, int fluid_frame, double *K576, double *CB144)
{
    double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
    double D[6][6], B[6][24], DB[6][24];
    dmat(lam, mu, D);
    memset(K576, 0, 576 * sizeof(double));
    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
        bmat(fluid_frame, GP[a], GP[b], GP[c], B);
        for (int i = 0; i < 6; i++) for (int j = 0; j < 24; j++) {
            double t = 0; for (int k = 0; k < 6; k++) t += D[i][k] * B[k][j];
            DB[i][j] = t;
        }
        for (int i = 0; i < 24; i++) for (int j = 0; j < 24; j++) {
            double t = 0; for (int k = 0; k < 6; k++) t += B[k][i] * DB[k][j];
            K576[i * 24 + j] += 0.125 * t;
        }
    }
    if (CB144) {
        bmat(fluid_frame, 0.5, 0.5, 0.5, B);
        for (int i = 0; i < 6; i++) for (int j = 0; j < 24; j++) {
            double t = 0; for (int k = 0; k < 6; k++) t += D[i][k] * B[k][j];
            CB144[i * 24 + j] = t;
        }
    }
}

/* Stokes 32x32 [K G; G^T S], K = int 2 eps:eps (mu = 1), G_(i,a),j = int dN_i/dx_a N_j,
 * S = tau int gradN_i . gradN_j with tau = 1/6 for a unit voxel, row-major.
 * CB80 = [0.125 velocity-averaging rows (3x24) | 0.125 body-force weights (8)]
 * (femhmg_fluid_3D.c:41-110). */
void orc_ke_fluid(double *K1024, double *CB80)
{
    double Kv[576];
    orc_ke_elastic(1.0, 0.0, 1, Kv, NULL); /* lambda = 0, mu = 1/2 ... */
    /* E = 1, nu = 0 gives mu = 1/2; the viscous form needs mu = 1: scale by 2 */
    memset(K1024, 0, 1024 * sizeof(double));
    for (int i = 0; i < 24; i++) for (int j = 0; j < 24; j++) K1024[i * 32 + j] = 2.0 * Kv[i * 24 + j];
    for (int a = 0; a < 2; a++) for (int b = 0; b < 2; b++) for (int c = 0; c < 2; c++) {
        double g[8][3], N[8];
        for (int k = 0; k < 8; k++) grad_shape(k, 1, GP[a], GP[b], GP[c], g[k], &N[k]);
        for (int i = 0; i < 8; i++) for (int d = 0; d < 3; d++) for (int j = 0; j < 8; j++) {
            double v = 0.125 * g[i][d] * N[j];
            K1024[(3 * i + d) * 32 + 24 + j] += v;
            K1024[(24 + j) * 32 + 3 * i + d] += v;
        }
        for (int i = 0; i < 8; i++) for (int j = 0; j < 8; j++)
            K1024[(24 + i) * 32 + 24 + j] += 0.125 * (1.0 / 6.0) *
                (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
    }
    if (CB80) {
        memset(CB80, 0, 80 * sizeof(double));
        for (int k = 0; k < 8; k++) for (int d = 0; d < 3; d++) CB80[d * 24 + 3 * k + d] = 0.125;
        for (int k = 0; k < 8; k++) CB80[72 + k] = 0.125;
    }
}

/* reference storage of the thermal matrix: the 40 non-zeros of each row, row
 * by row, ascending columns (femhmg_thermal_3D.c:42-120). */
void orc_thermal_pack40(const double *K64, double k, double *K40)
{
    int n = 0;
    for (int i = 0; i < 8; i++) for (int j = 0; j < 8; j++) {
        int diff = (DR[i] != DR[j]) + (DC[i] != DC[j]) + (DL[i] != DL[j]);
        if (diff != 1) K40[n++] = k * K64[i * 8 + j];
    }
}

/* =========================================================================
 * Model handle
 * ========================================================================= */
typedef struct {
    int analysis;
    unsigned nx, ny, nz, nmat;
    size_t N, ndof;
    double voxel;
    uint16_t *elem;      /* material index per element */
    uint16_t *keys;      /* thermal/elastic NBN keys */
    uint8_t *fkeys;      /* fluid node keys */
    uint32_t *dof_id;    /* fluid node -> dof id */
    uint32_t *dof2node;  /* fluid dof id -> node */
    uint8_t *bkeys;      /* fluid border keys */
    uint32_t nvel, nborder;
    size_t npore;
    double *K;           /* dense per material: 64 (thermal), 576 (elastic); 1024 (fluid) */
    double *CB;          /* 24 per mat thermal (3x8), 144 per mat elastic, 80 fluid */
    double props[512];
    double alpha[256];   /* thermal expansion coefficient per material (elastic) */
    double *esc;         /* scalar density field: per-element scale of K and CB (NULL: none) */
} orc_model;

static inline double esc_of(const orc_model *m, size_t e) { return m->esc ? m->esc[e] : 1.0; }

/* Scalar density field input (chfem/src/femhmg/femhmg.c:1760-1835): every
 * element's local matrices are those of its material times its own scale
 * (operator, Jacobi, RHS and C alike: femhmg_thermal_3D.c:265-325,558-583,
 * femhmg_elastic_3D.c:541-811,1382-1538).  esc = NULL removes it. */
void orc_set_element_scale(void *h, const double *esc)
{
    orc_model *m = (orc_model *)h;
    free(m->esc);
    m->esc = NULL;
    if (!esc) return;
    m->esc = (double *)malloc(m->N * sizeof(double));
    memcpy(m->esc, esc, m->N * sizeof(double));
}

void *orc_create(int analysis, unsigned nx, unsigned ny, unsigned nz, unsigned nmat,
                 const uint16_t *elem, const double *props, double voxel)
{
    orc_model *m = (orc_model *)calloc(1, sizeof(orc_model));
    m->analysis = analysis; m->nx = nx; m->ny = ny; m->nz = nz; m->nmat = nmat;
    m->N = (size_t)nx * ny * nz; m->voxel = voxel;
    memcpy(m->props, props, sizeof(double) * (analysis == ORC_ELASTIC ? 2 * nmat : nmat));
    m->elem = (uint16_t *)malloc(m->N * sizeof(uint16_t));
    memcpy(m->elem, elem, m->N * sizeof(uint16_t));
    if (analysis == ORC_THERMAL) {
        double K64[64];
        orc_ke_thermal_unit(K64);
        m->K = (double *)malloc(sizeof(double) * 64 * nmat);
        m->CB = (double *)malloc(sizeof(double) * 24 * nmat);
        for (unsigned t = 0; t < nmat; t++) {
            for (int i = 0; i < 64; i++) m->K[t * 64 + i] = props[t] * K64[i];
            /* flux at the element centre: k * grad N (femhmg_thermal_3D.c:432-493) */
            for (int k = 0; k < 8; k++) {
                double g[3];
                grad_shape(k, 0, 0.5, 0.5, 0.5, g, NULL);
                for (int d = 0; d < 3; d++) m->CB[t * 24 + d * 8 + k] = props[t] * g[d];
            }
        }
        m->keys = (uint16_t *)malloc(m->N * sizeof(uint16_t));
        orc_dof_material_map(nx, ny, nz, elem, m->keys);
        m->ndof = m->N;
    } else if (analysis == ORC_ELASTIC) {
        m->K = (double *)malloc(sizeof(double) * 576 * nmat);
        m->CB = (double *)malloc(sizeof(double) * 144 * nmat);
        for (unsigned t = 0; t < nmat; t++)
            orc_ke_elastic(props[2 * t], props[2 * t + 1], 0, m->K + 576 * t, m->CB + 144 * t);
        m->keys = (uint16_t *)malloc(m->N * sizeof(uint16_t));
        orc_dof_material_map(nx, ny, nz, elem, m->keys);
        m->ndof = 3 * m->N;
    } else {
        m->K = (double *)malloc(sizeof(double) * 1024);
        m->CB = (double *)malloc(sizeof(double) * 80);
        orc_ke_fluid(m->K, m->CB);
        m->fkeys = (uint8_t *)malloc(m->N);
        m->npore = orc_fluid_keys(nx, ny, nz, elem, m->fkeys);
        m->dof_id = (uint32_t *)malloc(m->N * sizeof(uint32_t));
        orc_fluid_dof_ids(m->N, m->fkeys, m->dof_id, &m->nvel, &m->nborder, NULL, NULL);
        m->dof2node = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)m->nvel + m->nborder + 1));
        m->bkeys = (uint8_t *)malloc((size_t)m->nborder + 1);
        orc_fluid_dof_ids(m->N, m->fkeys, m->dof_id, &m->nvel, &m->nborder, m->dof2node, m->bkeys);
        m->ndof = 4 * (size_t)m->nvel + m->nborder;
    }
    return m;
}

void orc_destroy(void *h)
{
    orc_model *m = (orc_model *)h;
    if (!m) return;
    free(m->elem); free(m->keys); free(m->fkeys); free(m->dof_id); free(m->dof2node);
    free(m->bkeys); free(m->K); free(m->CB); free(m->esc); free(m);
}

size_t orc_ndof(void *h) { return ((orc_model *)h)->ndof; }
uint32_t orc_nvel(void *h) { return ((orc_model *)h)->nvel; }
uint32_t orc_nborder(void *h) { return ((orc_model *)h)->nborder; }
size_t orc_npore(void *h) { return ((orc_model *)h)->npore; }
void orc_get_keys(void *h, uint16_t *out) { orc_model *m = h; memcpy(out, m->keys, m->N * 2); }
void orc_get_fluid_maps(void *h, uint8_t *fkeys, uint32_t *dof_id, uint32_t *dof2node, uint8_t *bkeys)
{
    orc_model *m = h;
    if (fkeys) memcpy(fkeys, m->fkeys, m->N);
    if (dof_id) memcpy(dof_id, m->dof_id, m->N * 4);
    if (dof2node) memcpy(dof2node, m->dof2node, ((size_t)m->nvel + m->nborder) * 4);
    if (bkeys) memcpy(bkeys, m->bkeys, m->nborder);
}
void orc_get_local(void *h, double *K, double *CB)
{
    orc_model *m = h;
    size_t kn = m->analysis == ORC_THERMAL ? 64 : (m->analysis == ORC_ELASTIC ? 576 : 1024);
    size_t cn = m->analysis == ORC_THERMAL ? 24 : (m->analysis == ORC_ELASTIC ? 144 : 80);
    size_t nm = m->analysis == ORC_FLUID ? 1 : m->nmat;
    if (K) memcpy(K, m->K, kn * nm * sizeof(double));
    if (CB) memcpy(CB, m->CB, cn * nm * sizeof(double));
}

/* =========================================================================
 * A7/A8: node-by-node operator q = A v.  For node i and each e in 0..7 the
 *     element for which i is local node e has material (key >> 2e) & 3; its
 *     rows for local node e multiply the element's nodal values
 *     (chfem/src/femhmg/cudapcg/kernels/matrix_free.cu:191-323 thermal SparseQ1,
 *     :416-459 elastic, kernels :1936-1954 and :2905-2925).
 * A10: Stokes pore and border rows (matrix_free.cu:674-732, :927-977).
 * ========================================================================= */
static void aprod_thermal_nbn(const orc_model *m, const double *v, double *q)
{
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)m->N; i++) {
        uint32_t nb[27];
        nbr27((size_t)i, ny, nx, nz, nb);
        unsigned key = m->keys[i];
        double res = 0.0;
        for (int e = 0; e < 8; e++) {
            const double *K = m->K + 64 * ((key >> (2 * e)) & 3u) + 8 * e;
            for (int j = 0; j < 8; j++)
                res += K[j] * v[nb[NB_SLOT(e, j)]];
        }
        q[i] = res;
    }
}

static void aprod_elastic_nbn(const orc_model *m, const double *v, double *q)
{
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)m->N; i++) {
        uint32_t nb[27];
        nbr27((size_t)i, ny, nx, nz, nb);
        unsigned key = m->keys[i];
        double rx = 0, ry = 0, rz = 0;
        for (int e = 0; e < 8; e++) {
            const double *K = m->K + 576 * ((key >> (2 * e)) & 3u) + 72 * e;
            for (int j = 0; j < 8; j++) {
                size_t d = 3 * (size_t)nb[NB_SLOT(e, j)];
                double vx = v[d], vy = v[d + 1], vz = v[d + 2];
                rx += K[3 * j] * vx + K[3 * j + 1] * vy + K[3 * j + 2] * vz;
                ry += K[24 + 3 * j] * vx + K[24 + 3 * j + 1] * vy + K[24 + 3 * j + 2] * vz;
                rz += K[48 + 3 * j] * vx + K[48 + 3 * j + 1] * vy + K[48 + 3 * j + 2] * vz;
            }
        }
        q[3 * i] = rx; q[3 * i + 1] = ry; q[3 * i + 2] = rz;
    }
}

/* element-by-element form: per element K_e u_e scattered to its 8 nodes
 * (matrix_free.cu:2187-2268 thermal, :3164-3262 elastic).  The scatter is done
 * node by node here (gathering the 8 element results) so the loop stays
 * race-free under OpenMP; the sum is the same set of terms. */
static void aprod_ebe(const orc_model *m, const double *v, double *q)
{
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    const int nd = m->analysis == ORC_ELASTIC ? 3 : 1, ke = 8 * nd, ksz = ke * ke;
    double *fe = (double *)malloc(sizeof(double) * m->N * ke);
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < (long)m->N; e++) {
        uint32_t nb[27];
        nbr27((size_t)e, ny, nx, nz, nb);
        double ue[24];
        for (int j = 0; j < 8; j++) {
            size_t d = (size_t)nd * nb[NB_ELEM(j)];
            for (int a = 0; a < nd; a++) ue[nd * j + a] = v[d + a];
        }
        const double *K = m->K + (size_t)ksz * m->elem[e];
        const double sc = esc_of(m, (size_t)e);
        for (int i = 0; i < ke; i++) {
            double t = 0; for (int j = 0; j < ke; j++) t += K[i * ke + j] * ue[j];
            fe[(size_t)e * ke + i] = sc * t;
        }
    }
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)m->N; i++) {
        uint32_t nb[27];
        nbr27((size_t)i, ny, nx, nz, nb);
        for (int a = 0; a < nd; a++) {
            double t = 0;
            for (int e = 0; e < 8; e++)
                t += fe[(size_t)nb[NB_ORIG(e)] * ke + nd * e + a];
            q[(size_t)nd * i + a] = t;
        }
    }
    free(fe);
}

static void aprod_fluid(const orc_model *m, const double *v, double *q)
{
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    const uint32_t nvel = m->nvel, ntot = m->nvel + m->nborder;
    const double *K = m->K;
    #pragma omp parallel for schedule(dynamic, 1024)
    for (long id = 0; id < (long)ntot; id++) {
        uint32_t i = m->dof2node[id];
        uint32_t nbi[27];
        nbr27((size_t)i, ny, nx, nz, nbi);
        int pore = id < (long)nvel;
        unsigned fkey = pore ? 255u : m->bkeys[id - nvel];
        double rx = 0, ry = 0, rz = 0, rp = 0;
        for (int e = 0; e < 8; e++) {
            if (!((fkey >> e) & 1u)) continue;
            for (int j = 0; j < 8; j++) {
                uint32_t nb = m->dof_id[nbi[NB_SLOT(e, j)]];
                if (nb < nvel) {
                    double vx = v[3 * (size_t)nb], vy = v[3 * (size_t)nb + 1], vz = v[3 * (size_t)nb + 2];
                    if (pore) {
                        rx += K[(3 * e) * 32 + 3 * j] * vx + K[(3 * e) * 32 + 3 * j + 1] * vy + K[(3 * e) * 32 + 3 * j + 2] * vz;
                        ry += K[(3 * e + 1) * 32 + 3 * j] * vx + K[(3 * e + 1) * 32 + 3 * j + 1] * vy + K[(3 * e + 1) * 32 + 3 * j + 2] * vz;
                        rz += K[(3 * e + 2) * 32 + 3 * j] * vx + K[(3 * e + 2) * 32 + 3 * j + 1] * vy + K[(3 * e + 2) * 32 + 3 * j + 2] * vz;
                    }
                    rp += K[(24 + e) * 32 + 3 * j] * vx + K[(24 + e) * 32 + 3 * j + 1] * vy + K[(24 + e) * 32 + 3 * j + 2] * vz;
                }
                double vp = v[3 * (size_t)nvel + nb];
                if (pore) {
                    rx += K[(3 * e) * 32 + 24 + j] * vp;
                    ry += K[(3 * e + 1) * 32 + 24 + j] * vp;
                    rz += K[(3 * e + 2) * 32 + 24 + j] * vp;
                }
                rp -= K[(24 + e) * 32 + 24 + j] * vp;
            }
        }
        if (pore) { q[3 * id] = rx; q[3 * id + 1] = ry; q[3 * id + 2] = rz; }
        q[3 * (size_t)nvel + id] = rp;
    }
}

/* q = A v; strategy 0 = NBN, 1 = EBE (thermal/elastic only) */
void orc_apply(void *h, int strategy, const double *v, double *q)
{
    orc_model *m = (orc_model *)h;
    if (m->analysis == ORC_FLUID) aprod_fluid(m, v, q);
    else if (strategy == 1 || m->esc) aprod_ebe(m, v, q);
    else if (m->analysis == ORC_THERMAL) aprod_thermal_nbn(m, v, q);
    else aprod_elastic_nbn(m, v, q);
}

/* =========================================================================
 * A11: Jacobi diagonal (inverse) recomputed from the keys
 *     (matrix_free.cu:167-188 thermal, :388-408 elastic, :663-672 / :912-925
 *     Stokes pore/border).  The Stokes "diagonal" carries the negative
 *     stabilisation entries of the pressure rows.
 * ========================================================================= */
void orc_jacobi_inv(void *h, double *dinv)
{
    orc_model *m = (orc_model *)h;
    if (m->analysis == ORC_FLUID) {
        const double *K = m->K;
        double dv = 0, dp = 0;
        for (int e = 0; e < 8; e++) { dv += K[(3 * e) * 32 + 3 * e]; dp -= K[(24 + e) * 32 + 24 + e]; }
        for (size_t id = 0; id < m->nvel; id++) {
            dinv[3 * id] = 1.0 / dv; dinv[3 * id + 1] = 1.0 / dv; dinv[3 * id + 2] = 1.0 / dv;
            dinv[3 * (size_t)m->nvel + id] = 1.0 / dp;
        }
        for (size_t b = 0; b < m->nborder; b++) {
            double d = 0;
            for (int e = 0; e < 8; e++) if ((m->bkeys[b] >> e) & 1u) d -= K[(24 + e) * 32 + 24 + e];
            dinv[4 * (size_t)m->nvel + b] = 1.0 / d;
        }
        return;
    }
    const int nd = m->analysis == ORC_ELASTIC ? 3 : 1, ke = 8 * nd, ksz = ke * ke;
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)m->N; i++) {
        unsigned key = m->keys[i];
        long r = i % ny, c = (i / ny) % nx, l = i / (nx * ny);
        for (int a = 0; a < nd; a++) {
            double d = 0;
            for (int e = 0; e < 8; e++) {
                double kd = m->K[(size_t)ksz * ((key >> (2 * e)) & 3u) + (nd * e + a) * ke + nd * e + a];
                if (m->esc) kd *= m->esc[pidx(r - DR[e], c - DC[e], l - DL[e], ny, nx, nz)];
                d += kd;
            }
            dinv[nd * i + a] = 1.0 / d;
        }
    }
}

/* =========================================================================
 * A6: right-hand sides.  Seam-jump formulation: the unknown is the total field
 *     with a jump of L (element units) across one periodic seam, so
 *     b = -sum over seam elements of K_e[:, seam columns] * L
 *     (femhmg_thermal_3D.c:122-236, femhmg_elastic_3D.c:130-462).
 *     Stokes: unit body force on every pore element, added to velocity rows
 *     only (femhmg_fluid_3D.c:112-181).
 * ========================================================================= */
typedef struct { int seam_axis; int S[4]; int comp; } loadcase_t;
/* seam_axis: 0 = last column (x), 1 = first row (y), 2 = first layer (z) */
static const loadcase_t LOADCASES[6] = {
    {0, {1, 2, 5, 6}, 0},  /* X  */
    {1, {2, 3, 6, 7}, 1},  /* Y  */
    {2, {0, 1, 2, 3}, 2},  /* Z  */
    {1, {2, 3, 6, 7}, 2},  /* YZ */
    {2, {0, 1, 2, 3}, 0},  /* XZ */
    {1, {2, 3, 6, 7}, 0},  /* XY */
};

static size_t seam_count(const orc_model *m, int axis)
{
    return axis == 0 ? (size_t)m->ny * m->nz : (axis == 1 ? (size_t)m->nx * m->nz : (size_t)m->nx * m->ny);
}
/* i-th seam element in the reference's loop order */
static size_t seam_elem(const orc_model *m, int axis, size_t i)
{
    size_t ny = m->ny, nx = m->nx, nxy = nx * ny;
    if (axis == 0) return (nx - 1) * ny + i % ny + (i / ny) * nxy;
    if (axis == 1) return (i % nx) * ny + (i / nx) * nxy;
    return i;
}
static double seam_L(const orc_model *m, int axis) { return axis == 0 ? m->nx : (axis == 1 ? m->ny : m->nz); }

static void rhs_thermal_expansion(const orc_model *m, double *b);

/* dir 0..5: seam-jump load cases; dir 6 (elastic): thermal expansion */
void orc_rhs(void *h, int dir, double *b)
{
    orc_model *m = (orc_model *)h;
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    memset(b, 0, m->ndof * sizeof(double));
    if (dir == ORC_DIR_TEXP) {
        rhs_thermal_expansion(m, b);
        return;
    }
    if (m->analysis == ORC_FLUID) {
        for (size_t e = 0; e < m->N; e++) {
            if (m->elem[e] != 0) continue;
            long r = e % ny, c = (e / ny) % nx, l = e / (nx * ny);
            for (int k = 0; k < 8; k++) {
                uint32_t id = m->dof_id[pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz)];
                if (id < m->nvel) b[3 * (size_t)id + dir] += m->CB[72 + k];
            }
        }
        return;
    }
    const int nd = m->analysis == ORC_ELASTIC ? 3 : 1, ke = 8 * nd, ksz = ke * ke;
    const loadcase_t lc = LOADCASES[dir];
    const int comp = nd == 3 ? lc.comp : 0;
    const double L = seam_L(m, lc.seam_axis);
    const size_t ns = seam_count(m, lc.seam_axis);
    for (size_t i = 0; i < ns; i++) {
        size_t e = seam_elem(m, lc.seam_axis, i);
        long r = e % ny, c = (e / ny) % nx, l = e / (nx * ny);
        const double *K = m->K + (size_t)ksz * m->elem[e];
        const double sc = esc_of(m, e);
        for (int k = 0; k < 8; k++) {
            size_t d = (size_t)nd * pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz);
            for (int a = 0; a < nd; a++) {
                double s = 0;
                for (int t = 0; t < 4; t++) s += K[(nd * k + a) * ke + nd * lc.S[t] + comp];
                b[d + a] -= sc * s * L;
            }
        }
    }
}

/* =========================================================================
 * A17: effective tensor column (femhmg_thermal_3D.c:355-500,
 *     femhmg_elastic_3D.c:1075-1342, femhmg_fluid_3D.c:184-285).
 *     C[row*dim + dir] = (seam term + sum_e CB_e u_e) / N; Stokes:
 *     voxel^2 / N * sum over pore elements of the averaged velocity.
 * ========================================================================= */
void orc_updateC(void *h, int dir, const double *x, double *C)
{
    orc_model *m = (orc_model *)h;
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    if (m->analysis == ORC_FLUID) {
        double acc[3] = {0, 0, 0};
        for (size_t e = 0; e < m->N; e++) {
            if (m->elem[e] != 0) continue;
            long r = e % ny, c = (e / ny) % nx, l = e / (nx * ny);
            double loc[3] = {0, 0, 0};
            for (int k = 0; k < 8; k++) {
                uint32_t id = m->dof_id[pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz)];
                if (id < m->nvel)
                    for (int a = 0; a < 3; a++) loc[a] += m->CB[a * 24 + 3 * k + a] * x[3 * (size_t)id + a];
            }
            for (int a = 0; a < 3; a++) acc[a] += loc[a];
        }
        for (int a = 0; a < 3; a++) C[a * 3 + dir] = acc[a] / (double)m->N * m->voxel * m->voxel;
        return;
    }
    const int nd = m->analysis == ORC_ELASTIC ? 3 : 1, ke = 8 * nd, nrow = nd == 3 ? 6 : 3;
    const int cbsz = nrow * ke;
    const loadcase_t lc = LOADCASES[dir];
    const int comp = nd == 3 ? lc.comp : 0;
    const double L = seam_L(m, lc.seam_axis);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const size_t ns = seam_count(m, lc.seam_axis);
    for (size_t i = 0; i < ns; i++) {
        const size_t se = seam_elem(m, lc.seam_axis, i);
        const double *CB = m->CB + (size_t)cbsz * m->elem[se];
        const double sc = esc_of(m, se);
        for (int row = 0; row < nrow; row++)
            for (int t = 0; t < 4; t++) acc[row] += sc * CB[row * ke + nd * lc.S[t] + comp] * L;
    }
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0, a5 = 0;
    #pragma omp parallel for reduction(+ : a0, a1, a2, a3, a4, a5) schedule(static)
    for (long e = 0; e < (long)m->N; e++) {
        long r = e % ny, c = (e / ny) % nx, l = e / (nx * ny);
        const double *CB = m->CB + (size_t)cbsz * m->elem[e];
        const double sc = esc_of(m, (size_t)e);
        double loc[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 8; k++) {
            size_t d = (size_t)nd * pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz);
            for (int a = 0; a < nd; a++)
                for (int row = 0; row < nrow; row++) loc[row] += CB[row * ke + nd * k + a] * x[d + a];
        }
        a0 += sc * loc[0]; a1 += sc * loc[1]; a2 += sc * loc[2]; a3 += sc * loc[3]; a4 += sc * loc[4];
        a5 += sc * loc[5];
    }
    acc[0] += a0; acc[1] += a1; acc[2] += a2; acc[3] += a3; acc[4] += a4; acc[5] += a5;
    for (int row = 0; row < nrow; row++) C[row * nrow + dir] = acc[row] / (double)m->N;
}

/* =========================================================================
 * Thermal expansion (elastic only; femhmg_elastic_3D.c:464-508, 863-1073,
 * femhmg.c:1121-1165).  Every element carries the isotropic eigenstrain
 * a_e (1,1,1,0,0,0) for a unit temperature rise.  Load: f_e = B(centre)^T D_e
 * a_e scattered to the element's nodes, no seam jump.  With x the periodic
 * solution, sigma = <D_e (B x_e - a_e)> and eps = <B x_e> (averages over
 * elements, centre point), beta = -sigma + C eps and alpha* = C^-1 beta by
 * Gaussian elimination without pivoting, as femhmg_elastic_3D.c:1045-1063.
 * ========================================================================= */
void orc_set_alpha(void *h, const double *alpha)
{
    orc_model *m = (orc_model *)h;
    memcpy(m->alpha, alpha, sizeof(double) * m->nmat);
}

static void elastic_D(const orc_model *m, unsigned t, double D[6][6])
{
    const double E = m->props[2 * t], nu = m->props[2 * t + 1];
    dmat(E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu)), D);
}

static void rhs_thermal_expansion(const orc_model *m, double *b)
{
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    double B[6][24];
    bmat(0, 0.5, 0.5, 0.5, B);
    for (size_t e = 0; e < m->N; e++) {
        const unsigned t = m->elem[e];
        double D[6][6], Da[6];
        elastic_D(m, t, D);
        for (int i = 0; i < 6; i++) Da[i] = (D[i][0] + D[i][1] + D[i][2]) * m->alpha[t];
        long r = e % ny, c = (e / ny) % nx, l = e / (nx * ny);
        for (int k = 0; k < 8; k++) {
            size_t n = 3 * (size_t)pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz);
            for (int j = 0; j < 3; j++) {
                double f = 0;
                for (int i = 0; i < 6; i++) f += B[i][3 * k + j] * Da[i];
                b[n + j] += f;
            }
        }
    }
}

/* C: the 6x6 effective matrix (row-major); out: alpha* in Voigt order
 * xx, yy, zz, yz, xz, xy */
void orc_thermal_expansion(void *h, const double *x, const double *C, double *out)
{
    orc_model *m = (orc_model *)h;
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    double B[6][24], stress[6] = {0, 0, 0, 0, 0, 0}, strain[6] = {0, 0, 0, 0, 0, 0};
    bmat(0, 0.5, 0.5, 0.5, B);
    for (size_t e = 0; e < m->N; e++) {
        const unsigned t = m->elem[e];
        const double *CB = m->CB + 144 * (size_t)t;
        long r = e % ny, c = (e / ny) % nx, l = e / (nx * ny);
        double ue[24];
        for (int k = 0; k < 8; k++) {
            size_t n = 3 * (size_t)pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz);
            for (int j = 0; j < 3; j++) ue[3 * k + j] = x[n + j];
        }
        double D[6][6];
        elastic_D(m, t, D);
        for (int i = 0; i < 6; i++) {
            double s = 0, g = 0;
            for (int j = 0; j < 24; j++) { s += CB[i * 24 + j] * ue[j]; g += B[i][j] * ue[j]; }
            stress[i] += s - (D[i][0] + D[i][1] + D[i][2]) * m->alpha[t];
            strain[i] += g;
        }
    }
    double beta[6], M[36];
    for (int i = 0; i < 6; i++) { stress[i] /= (double)m->N; strain[i] /= (double)m->N; }
    for (int i = 0; i < 6; i++) {
        beta[i] = -stress[i];
        for (int j = 0; j < 6; j++) beta[i] += C[6 * i + j] * strain[j];
    }
    memcpy(M, C, sizeof(M));
    for (int i = 0; i < 5; i++)
        for (int j = i + 1; j < 6; j++) {
            const double f = M[6 * j + i] / M[7 * i];
            for (int k = i + 1; k < 6; k++) M[6 * j + k] -= M[6 * i + k] * f;
            beta[j] -= beta[i] * f;
            M[6 * j + i] = 0;
        }
    for (int i = 5; i >= 0; i--) {
        double a = beta[i];
        for (int j = 5; j > i; j--) a -= M[6 * i + j] * out[j];
        out[i] = a / M[7 * i];
    }
}

/* =========================================================================
 * Nodal fields for export (femhmg_thermal_3D.c:694-1105,
 * femhmg_elastic_3D.c:1800-2645, femhmg_fluid_3D.c:312-360).
 *   thermal: per node, 1/8 of k grad T (3 values) of every element that has
 *            the node as a corner, evaluated at that corner;
 *   elastic: the same with the stress D B u (6 values, Voigt order);
 *   then, for the seam elements of load case dir, the field of the seam
 *   jump: corners off the jumped face take the field of the jump L on the
 *   jumped nodes, corners on it the field of -L on the other nodes
 *   (the reference's split; the two differ by a rigid translation).
 * out: 3 (thermal) or 6 (elastic) values per periodic node.
 * ========================================================================= */
static void corner_field(const orc_model *m, unsigned t, int k, const double *ue, double *f)
{
    double xk, yk, zk;
    node_xyz(k, 0, &xk, &yk, &zk);
    if (m->analysis == ORC_THERMAL) {
        f[0] = f[1] = f[2] = 0;
        for (int j = 0; j < 8; j++) {
            double g[3];
            grad_shape(j, 0, xk, yk, zk, g, NULL);
            for (int a = 0; a < 3; a++) f[a] += g[a] * ue[j];
        }
        for (int a = 0; a < 3; a++) f[a] *= m->props[t];
        return;
    }
    double B[6][24], D[6][6], eps[6];
    bmat(0, xk, yk, zk, B);
    elastic_D(m, t, D);
    for (int i = 0; i < 6; i++) {
        eps[i] = 0;
        for (int j = 0; j < 24; j++) eps[i] += B[i][j] * ue[j];
    }
    for (int i = 0; i < 6; i++) {
        f[i] = 0;
        for (int j = 0; j < 6; j++) f[i] += D[i][j] * eps[j];
    }
}

void orc_fields(void *h, int dir, const double *x, double *out)
{
    orc_model *m = (orc_model *)h;
    const long ny = m->ny, nx = m->nx, nz = m->nz;
    const int nd = m->analysis == ORC_ELASTIC ? 3 : 1, nf = nd == 3 ? 6 : 3;
    memset(out, 0, sizeof(double) * nf * m->N);
    for (size_t e = 0; e < m->N; e++) {
        long r = e % ny, c = (e / ny) % nx, l = e / (nx * ny);
        size_t node[8];
        double ue[24], f[6];
        for (int k = 0; k < 8; k++) {
            node[k] = pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz);
            for (int a = 0; a < nd; a++) ue[nd * k + a] = x[nd * node[k] + a];
        }
        for (int k = 0; k < 8; k++) {
            corner_field(m, m->elem[e], k, ue, f);
            for (int i = 0; i < nf; i++) out[nf * node[k] + i] += 0.125 * f[i];
        }
    }
    if (dir < 0 || dir > 5) return;
    const loadcase_t lc = LOADCASES[dir];
    const int comp = nd == 3 ? lc.comp : 0;
    const double L = seam_L(m, lc.seam_axis);
    const size_t ns = seam_count(m, lc.seam_axis);
    for (size_t i = 0; i < ns; i++) {
        const size_t e = seam_elem(m, lc.seam_axis, i);
        long r = e % ny, c = (e / ny) % nx, l = e / (nx * ny);
        int on[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = 0; t < 4; t++) on[lc.S[t]] = 1;
        double jp[24], jm[24], f[6];
        memset(jp, 0, sizeof(jp));
        memset(jm, 0, sizeof(jm));
        for (int k = 0; k < 8; k++) {
            if (on[k]) jp[nd * k + comp] = L;
            else jm[nd * k + comp] = -L;
        }
        for (int k = 0; k < 8; k++) {
            const size_t n = pidx(r + DR[k], c + DC[k], l + DL[k], ny, nx, nz);
            corner_field(m, m->elem[e], k, on[k] ? jm : jp, f);
            for (int q = 0; q < nf; q++) out[nf * n + q] += 0.125 * f[q];
        }
    }
}

/* Stokes: velocity (3 per node, pore nodes only) and pressure (1 per node,
 * every node with a fluid key) from the DOF vector */
void orc_fluid_fields(void *h, const double *x, double *vel, double *pres)
{
    orc_model *m = (orc_model *)h;
    for (size_t n = 0; n < m->N; n++) {
        const uint32_t id = m->dof_id[n];
        vel[3 * n] = vel[3 * n + 1] = vel[3 * n + 2] = 0;
        pres[n] = 0;
        if (m->fkeys[n] == 255)
            for (int a = 0; a < 3; a++) vel[3 * n + a] = x[3 * (size_t)id + a];
        if (m->fkeys[n]) pres[n] = x[3 * (size_t)m->nvel + id];
    }
}

/* =========================================================================
 * A14/A15: Jacobi-preconditioned CG and MINRES with the reference's iteration
 *     order and L2 stopping rule (chfem/src/femhmg/cudapcg/solvers/pcg.cu:28-243,
 *     pminres.cu:28-247, stopping_criteria.cu:7-14).
 * ========================================================================= */
static double dot(const double *a, const double *b, size_t n)
{
    double s = 0;
    #pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < (long)n; i++) s += a[i] * b[i];
    return s;
}

static void precond(const double *dinv, int on, const double *v, double *out, size_t n)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) out[i] = on ? dinv[i] * v[i] : v[i];
}

static int above_tol(double delta, double delta0, double tol) { return fabs(delta) > tol * tol * fabs(delta0); }

static double absmax_n(const double *v, size_t n)
{
    double m = 0.0;
    for (size_t i = 0; i < n; i++) m = fmax(m, fabs(v[i]));
    return m;
}

/* The stopping rules (solvers/pcg.cu:130-146 and the switch at :176-190,
 * pminres.cu:188-200, stopping_criteria.cu:7-14):
 *   0 L2:    r.M^-1 r <= tol^2 r0.M^-1 r0
 *   1 INF:   max|r| / max|b| <= tol (b: the residual of the zero guess)
 *   2 error: |alpha| max|d| / max|x| <= tol over the first nhmg entries;
 *            CG also stops when the L2 rule is met
 * resnorm L2 is orc_solve.  Returns 1 if the rule was met. */
int orc_solve_rule(void *h, int dir, int solver, int precon, int strategy, double tol, unsigned maxit, int resnorm,
                   size_t nhmg, const double *x0, double *x, unsigned *iters, double *res_out)
{
    orc_model *m = (orc_model *)h;
    const size_t n = m->ndof;
    double *r = malloc(n * sizeof(double)), *d = malloc(n * sizeof(double));
    double *q = malloc(n * sizeof(double)), *s = malloc(n * sizeof(double));
    double *dinv = malloc(n * sizeof(double));
    orc_jacobi_inv(h, dinv);
    orc_rhs(h, dir, r);
    if (x0) memcpy(x, x0, n * sizeof(double)); else memset(x, 0, n * sizeof(double));
    memset(q, 0, n * sizeof(double));
    unsigned it = 0;
    double residual = 0;
    int found = 0;
    if (nhmg == 0 || nhmg > n) nhmg = n;
    const double rinf0 = absmax_n(r, n);
    double metric = 1.0;
    if (solver == 0) {
        precond(dinv, precon, r, d, n);
        double delta0 = dot(r, d, n), delta, delta_old, a;
        const double res0 = resnorm == 1 ? rinf0 : delta0;
        if (fabs(res0) < 1e-12) { memset(x, 0, n * sizeof(double)); found = 1; goto done; }
        if (x0) {
            orc_apply(h, strategy, x, q);
            for (size_t i = 0; i < n; i++) r[i] -= q[i];
            precond(dinv, precon, r, d, n);
            delta = dot(r, d, n);
            if (!above_tol(delta, delta0, tol)) { residual = sqrt(fabs(delta) / fabs(delta0)); found = 1; goto done; }
        } else delta = delta0;
        delta_old = delta;
        it = 1;
        orc_apply(h, strategy, d, q);
        a = delta / dot(d, q, n);
        #pragma omp parallel for
        for (long i = 0; i < (long)n; i++) { x[i] += a * d[i]; r[i] -= a * q[i]; }
        precond(dinv, precon, r, q, n);
        delta = dot(r, q, n);
        /* the rule after an iteration (a, d, x, r of that iteration) */
#define PCG_RULE()                                                                                 \
        (resnorm == 1 ? ((metric = absmax_n(r, n) / res0) > tol)                                   \
         : resnorm == 2 ? (metric = fabs(a) * absmax_n(d, nhmg) / absmax_n(x, nhmg),               \
                           !(metric <= tol || !above_tol(delta, delta0, tol)))                     \
                        : above_tol(delta, delta0, tol))
        int cont = PCG_RULE();
        while (cont && it < maxit) {
            it++;
            double beta = delta / delta_old;
            #pragma omp parallel for
            for (long i = 0; i < (long)n; i++) d[i] = q[i] + beta * d[i];
            orc_apply(h, strategy, d, q);
            a = delta / dot(d, q, n);
            #pragma omp parallel for
            for (long i = 0; i < (long)n; i++) { x[i] += a * d[i]; r[i] -= a * q[i]; }
            precond(dinv, precon, r, q, n);
            delta_old = delta;
            delta = dot(r, q, n);
            cont = PCG_RULE();
        }
#undef PCG_RULE
        if (resnorm == 0) {
            residual = sqrt(fabs(delta) / fabs(delta0));
            found = !above_tol(delta, delta0, tol);
        } else {
            residual = metric;
            found = metric <= tol;
        }
    } else {
        /* MINRES */
        double res0, delta0, delta, delta_old, a, b, stop;
        precond(dinv, precon, r, q, n);
        res0 = resnorm == 1 ? rinf0 : dot(r, q, n);
        orc_apply(h, strategy, q, s);
        delta0 = dot(q, s, n);
        if (fabs(res0) < 1e-12) { memset(x, 0, n * sizeof(double)); found = 1; goto done; }
        if (x0) {
            orc_apply(h, strategy, x, d);
            for (size_t i = 0; i < n; i++) r[i] -= d[i];
            precond(dinv, precon, r, q, n);
            orc_apply(h, strategy, q, s);
            delta = dot(q, s, n);
            if (!above_tol(delta, delta0, tol)) { residual = sqrt(fabs(delta) / fabs(delta0)); found = 1; goto done; }
        } else delta = delta0;
        delta_old = delta;
        it = 1;
        memcpy(d, q, n * sizeof(double));
        memcpy(q, s, n * sizeof(double));
        precond(dinv, precon, q, s, n);        /* s used as scratch for M^-1 q */
        a = delta / dot(q, s, n);
        #pragma omp parallel for
        for (long i = 0; i < (long)n; i++) { x[i] += a * d[i]; r[i] -= a * q[i]; }
        precond(dinv, precon, r, r, n);
        orc_apply(h, strategy, r, s);
        delta_old = delta;
        delta = dot(r, s, n);
        #pragma omp parallel for
        for (long i = 0; i < (long)n; i++) r[i] = precon ? r[i] / dinv[i] : r[i];
        {
            double *t = malloc(n * sizeof(double));
            precond(dinv, precon, r, t, n);
            stop = dot(r, t, n);
#define MINRES_RULE()                                                                              \
            (resnorm == 1 ? ((metric = absmax_n(r, n) / res0) > tol)                               \
             : resnorm == 2 ? ((metric = fabs(a) * absmax_n(d, nhmg) / absmax_n(x, nhmg)) > tol)   \
                            : above_tol(stop, res0, tol))
            int cont = MINRES_RULE();
            while (cont && it < maxit) {
                it++;
                b = delta / delta_old;
                #pragma omp parallel for
                for (long i = 0; i < (long)n; i++) {
                    d[i] = (precon ? dinv[i] * r[i] : r[i]) + b * d[i];
                    q[i] = s[i] + b * q[i];
                }
                precond(dinv, precon, q, t, n);
                a = delta / dot(q, t, n);
                #pragma omp parallel for
                for (long i = 0; i < (long)n; i++) { x[i] += a * d[i]; r[i] -= a * q[i]; }
                precond(dinv, precon, r, r, n);
                orc_apply(h, strategy, r, s);
                delta_old = delta;
                delta = dot(r, s, n);
                #pragma omp parallel for
                for (long i = 0; i < (long)n; i++) r[i] = precon ? r[i] / dinv[i] : r[i];
                precond(dinv, precon, r, t, n);
                stop = dot(r, t, n);
                cont = MINRES_RULE();
            }
#undef MINRES_RULE
            free(t);
        }
        residual = resnorm == 0 ? sqrt(fabs(stop) / fabs(res0)) : metric;
        found = residual <= tol;
    }
done:
    if (iters) *iters = it;
    if (res_out) *res_out = residual;
    free(r); free(d); free(q); free(s); free(dinv);
    return found;
}

/* returns 1 if converged; x holds the solution; iters/res reported */
int orc_solve(void *h, int dir, int solver, int precon, int strategy, double tol, unsigned maxit,
              const double *x0, double *x, unsigned *iters, double *res_out)
{
    return orc_solve_rule(h, dir, solver, precon, strategy, tol, maxit, 0, 0, x0, x, iters, res_out);
}

/* =========================================================================
 * A16: coarse -> fine initial guess: x2 projection of even nodes, then linear
 *     interpolation of odd rows, odd columns and odd layers with periodic
 *     neighbours (chfem/src/femhmg/cudapcg/kernels/image.cu:19-121,
 *     wrappers.cu:290-303).  rows/cols/layers are the FINE element counts.
 *     Also the coarse material map: every second element in each direction
 *     (femhmg.c:1844-1916).
 * ========================================================================= */
void orc_interpl2(const double *coarse, unsigned rows, unsigned cols, unsigned layers, unsigned stride, double *fine)
{
    const size_t ny = rows, nx = cols, nz = layers, nxy = nx * ny, N = nxy * nz;
    const size_t nyc = ny / 2, nxc = nx / 2, nzc = nz / 2;
    memset(fine, 0, N * stride * sizeof(double));
    for (size_t i = 0; i < nxc * nyc * (nzc + (nzc < 1)); i++) {
        size_t ii = i % nyc, jj = (i % (nxc * nyc)) / nyc, kk = i / (nxc * nyc);
        size_t idx = 2 * ii + 2 * jj * ny + 2 * kk * nxy;
        for (unsigned j = 0; j < stride; j++) fine[stride * idx + j] = 2.0 * coarse[stride * i + j];
    }
    for (size_t i = 0; i < N; i++) {
        size_t ii = i % ny, jj = (i % nxy) / ny, kk = i / nxy;
        if ((ii & 1) && !(jj & 1) && !(kk & 1)) {
            size_t up = ii == 0 ? i - 1 + ny : i - 1;
            size_t dn = (ii + 1) == ny ? i + 1 - ny : i + 1;
            for (unsigned j = 0; j < stride; j++)
                fine[stride * i + j] = 0.5 * fine[stride * up + j] + 0.5 * fine[stride * dn + j];
        }
    }
    for (size_t i = 0; i < N; i++) {
        size_t jj = (i % nxy) / ny, kk = i / nxy;
        if ((jj & 1) && !(kk & 1)) {
            size_t loc = i % nxy;
            size_t left = (loc + nxy - ny) % nxy + kk * nxy, right = (loc + ny) % nxy + kk * nxy;
            for (unsigned j = 0; j < stride; j++)
                fine[stride * i + j] = 0.5 * fine[stride * left + j] + 0.5 * fine[stride * right + j];
        }
    }
    for (size_t i = 0; i < N; i++) {
        size_t kk = i / nxy;
        if (kk & 1) {
            size_t near = (i + N - nxy) % N, far = (i + nxy) % N;
            for (unsigned j = 0; j < stride; j++)
                fine[stride * i + j] = 0.5 * fine[stride * near + j] + 0.5 * fine[stride * far + j];
        }
    }
}

void orc_coarse_elem_map(const uint16_t *elem, unsigned nx, unsigned ny, unsigned nz, uint16_t *coarse)
{
    unsigned nxc = nx / 2, nyc = ny / 2, nzc = nz / 2;
    for (unsigned i = 0; i < nyc; i++)
        for (unsigned j = 0; j < nxc; j++)
            for (unsigned k = 0; k < nzc; k++)
                coarse[i + j * nyc + (size_t)k * nxc * nyc] = elem[2 * i + 2 * j * ny + (size_t)2 * k * nx * ny];
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }
double orc_wtime(void) { return omp_get_wtime(); }
