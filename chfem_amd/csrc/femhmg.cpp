// femhmg.cpp -- the homogenization-model C ABI (include/femhmg.h): neutral-file
// and image input, node keys / Stokes DOF maps, local matrices, the
// per-direction solve loop and the effective tensor.  Right-hand sides and
// effective-tensor columns are computed on the device; the solution vector
// stays in HBM unless it is exported.
//
// Reference behaviour followed: chfem/src/femhmg/femhmg.c:135-1198 (driver),
// :1456-1758 (input), chfem/src/femhmg/physical_phenomena/femhmg_3D.c,
// femhmg_thermal_3D.c, femhmg_elastic_3D.c, femhmg_fluid_3D.c (physics),
// chfem/src/main.c:72-152, :300-323 (runAnalysis / defaults).
#include <hip/hip_runtime.h>
#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "batch.h"
#include "noalloc.h"
#include "guess.h"
#include "solver.h"
#include "voxel.h"
#include "../../include/femhmg.h"
#include "../../include/chfem_amd.h"

using namespace chfem;

namespace {

// local node k of element (r,c,l) sits at (r+DR[k], c+DC[k], l+DL[k])
const int DR[8] = {1, 1, 0, 0, 1, 1, 0, 0};
const int DC[8] = {0, 1, 1, 0, 0, 1, 1, 0};
const int DL[8] = {0, 0, 0, 0, 1, 1, 1, 1};

struct Model {
    std::string nf, nf_noext, image_file, x0_file;
    bool have_sdf = false;            // scalar density field read from a .bin file
    bool sdf_f64 = false;             // %data_type float64 (else float32)
    std::vector<uint16_t> density;    // quantised field per (refined) element
    double dmin = 0, dmax = 0;
    int analysis = HMG_THERMAL;
    int dim = HMG_3D;
    unsigned hmg_flag = HOMOGENIZE_ALL;
    bool hmg_flag_set = false, using_x0 = false, export_x = false, thermal_expansion = false;
    bool save_fields = false;
    cudapcgFlag_t poremap = CUDAPCG_POREMAP_NUM;
    cudapcgFlag_t pcg_flag = CUDAPCG_THERMAL_3D;
    double elem_size = 1.0;
    unsigned refinement = 1;
    unsigned C_dim = 9, lclMtx_dim = 40, lclCB_dim = 0;
    double props[512];
    double alpha[256];
    double texp[6];                   // effective thermal expansion (Voigt xx,yy,zz,yz,xz,xy)
    bool texp_ok = false;
    uint8_t props_keys[256];
    unsigned nmat = 0;
    unsigned ex = 0, ey = 0, ez = 0;   // element counts (cols, rows, layers)
    size_t nelem = 0, ndof = 0;
    unsigned nvel = 0, nborder = 0;
    size_t nporeelem = 0;
    double tol = 1e-5;
    unsigned maxit = 10000;
    std::vector<uint16_t> elem_map, keys;
    std::vector<uint8_t> fkeys, bkeys;
    std::vector<uint32_t> dof_id, dof2node;
    std::vector<double> Mtxs, CB;        // reference layouts
    double C[36];
    std::vector<std::vector<double>> x0;  // per direction (recursive guesses)
    std::vector<chfemSolveStats_t> stats;
    std::vector<std::string> report;
    bool report_on = false;
};

Model *M = nullptr;

// driver flags (femhmg.c:23-25)
cudapcgFlag_t g_strategy = CUDAPCG_NBN, g_solver = CUDAPCG_CG_SOLVER, g_precond = CUDAPCG_TRUE;
cudapcgFlag_t g_poremap_default = CUDAPCG_POREMAP_NUM, g_xreduce = CUDAPCG_XREDUCE_FULL, g_stopcrit = CUDAPCG_L2_NORM;
cudapcgFlag_t g_precision = CHFEM_PREC_FP64;
double g_xreduce_stab = 0.0001;
double g_time_total = 0;

// ---------------------------------------------------------------------------
// local matrices of the unit voxel (voxel.h)
// ---------------------------------------------------------------------------
using voxel::elastic_K;
using voxel::frame;
using voxel::int_dd;
using voxel::lap;

// 6x24 centre stress operator, rows xx,yy,zz,yz,xz,xy (engineering shear)
void elastic_CB(double lam, double mu, double *CB)
{
    for (int k = 0; k < 8; k++) {
        int x, y, z;
        frame(k, false, x, y, z);
        const double g[3] = {(x ? 0.25 : -0.25), (y ? 0.25 : -0.25), (z ? 0.25 : -0.25)};
        for (int b = 0; b < 3; b++) {
            double eps[6] = {0, 0, 0, 0, 0, 0};
            eps[b] = g[b];
            if (b == 1) { eps[3] += g[2]; eps[5] += g[0]; }
            if (b == 2) { eps[3] += g[1]; eps[4] += g[0]; }
            if (b == 0) { eps[4] += g[2]; eps[5] += g[1]; }
            const double tr = eps[0] + eps[1] + eps[2];
            for (int r = 0; r < 3; r++) CB[r * 24 + 3 * k + b] = lam * tr + 2 * mu * eps[r];
            for (int r = 3; r < 6; r++) CB[r * 24 + 3 * k + b] = mu * eps[r];
        }
    }
}

void assemble_local(Model &m)
{
    if (m.analysis == HMG_THERMAL) {
        m.Mtxs.assign(40 * m.nmat, 0.0);
        for (unsigned t = 0; t < m.nmat; t++) {
            int n = 0;
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++) {
                    const int diff = (DR[i] != DR[j]) + (DC[i] != DC[j]) + (DL[i] != DL[j]);
                    if (diff != 1) m.Mtxs[40 * t + n++] = m.props[t] * lap(i, j, false);
                }
        }
        // thermal flux at the centre, 3x8 per material: k * grad N
        m.CB.assign(24 * m.nmat, 0.0);
        for (unsigned t = 0; t < m.nmat; t++)
            for (int k = 0; k < 8; k++) {
                int x, y, z;
                frame(k, false, x, y, z);
                m.CB[24 * t + 0 * 8 + k] = m.props[t] * (x ? 0.25 : -0.25);
                m.CB[24 * t + 1 * 8 + k] = m.props[t] * (y ? 0.25 : -0.25);
                m.CB[24 * t + 2 * 8 + k] = m.props[t] * (z ? 0.25 : -0.25);
            }
    } else if (m.analysis == HMG_ELASTIC) {
        m.Mtxs.assign(576 * m.nmat, 0.0);
        m.CB.assign(144 * m.nmat, 0.0);
        for (unsigned t = 0; t < m.nmat; t++) {
            const double E = m.props[2 * t], nu = m.props[2 * t + 1];
            const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
            elastic_K(lam, mu, false, &m.Mtxs[576 * t]);
            elastic_CB(lam, mu, &m.CB[144 * t]);
        }
    } else {
        // Stokes [K G; G^T S]: K = 2 eps:eps (mu = 1), G = int dN_i/dx_a N_j,
        // S = tau lap with tau = 1/6 (femhmg_fluid_3D.c:41-110)
        m.Mtxs.assign(1024, 0.0);
        std::vector<double> Kv(576);
        elastic_K(0.0, 1.0, true, Kv.data());
        for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) m.Mtxs[i * 32 + j] = Kv[i * 24 + j];
        for (int i = 0; i < 8; i++)
            for (int a = 0; a < 3; a++)
                for (int j = 0; j < 8; j++) {
                    const double gv = int_dd(i, j, a, -1, true);
                    m.Mtxs[(3 * i + a) * 32 + 24 + j] = gv;
                    m.Mtxs[(24 + j) * 32 + 3 * i + a] = gv;
                }
        for (int i = 0; i < 8; i++)
            for (int j = 0; j < 8; j++) m.Mtxs[(24 + i) * 32 + 24 + j] = lap(i, j, true) / 6.0;
        m.CB.assign(80, 0.0);
        for (int k = 0; k < 8; k++)
            for (int a = 0; a < 3; a++) m.CB[a * 24 + 3 * k + a] = 0.125;
        for (int k = 0; k < 8; k++) m.CB[72 + k] = 0.125;
    }
}

// dense per-material local matrix (thermal expanded from the 40 stored entries)
std::vector<double> dense_K(const Model &m)
{
    if (m.analysis == HMG_THERMAL) {
        std::vector<double> K(64 * m.nmat, 0.0);
        for (unsigned t = 0; t < m.nmat; t++) {
            int n = 0;
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++) {
                    const int diff = (DR[i] != DR[j]) + (DC[i] != DC[j]) + (DL[i] != DL[j]);
                    K[64 * t + i * 8 + j] = diff == 1 ? 0.0 : m.Mtxs[40 * t + n++];
                }
        }
        return K;
    }
    return m.Mtxs;
}

// ---------------------------------------------------------------------------
// maps (femhmg_3D.c:4-75, femhmg_fluid_3D.c:370-491), gathered node by node
// ---------------------------------------------------------------------------
inline size_t pid(long r, long c, long l, long ny, long nx, long nz)
{
    r = (r + ny) % ny; c = (c + nx) % nx; l = (l + nz) % nz;
    return (size_t)r + (size_t)c * ny + (size_t)l * nx * ny;
}

void build_keys(Model &m)
{
    const long ny = m.ey, nx = m.ex, nz = m.ez;
    m.keys.assign(m.nelem, 0);
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)m.nelem; i++) {
        const long r = i % ny, c = (i / ny) % nx, l = i / (nx * ny);
        unsigned key = 0;
        for (int e = 0; e < 8; e++)
            key |= (unsigned)(m.elem_map[pid(r - DR[e], c - DC[e], l - DL[e], ny, nx, nz)] & MATKEY_BITSTEP_RANGE_3D)
                   << (MATKEY_BITSTEP_3D * e);
        m.keys[i] = (uint16_t)key;
    }
}

void build_fluid_maps(Model &m)
{
    const long ny = m.ey, nx = m.ex, nz = m.ez;
    m.fkeys.assign(m.nelem, 0);
    size_t npore = 0;
    #pragma omp parallel for schedule(static) reduction(+ : npore)
    for (long i = 0; i < (long)m.nelem; i++) {
        const long r = i % ny, c = (i / ny) % nx, l = i / (nx * ny);
        unsigned key = 0;
        for (int e = 0; e < 8; e++)
            key |= (unsigned)(m.elem_map[pid(r - DR[e], c - DC[e], l - DL[e], ny, nx, nz)] == 0) << e;
        m.fkeys[i] = (uint8_t)key;
        npore += m.elem_map[i] == 0;
    }
    m.nporeelem = npore;
    // DOF numbering: nodes inside the pore space first, then interface nodes,
    // each in natural order; solid nodes map to 0 (femhmg_fluid_3D.c:370-424)
    m.dof_id.assign(m.nelem, 0);
    uint32_t count = 0;
    for (size_t i = 0; i < m.nelem; i++)
        if (m.fkeys[i] == 255) m.dof_id[i] = count++;
    m.nvel = count;
    for (size_t i = 0; i < m.nelem; i++)
        if (m.fkeys[i] && m.fkeys[i] != 255) m.dof_id[i] = count++;
    m.nborder = count - m.nvel;
    m.ndof = 4 * (size_t)m.nvel + m.nborder;
    m.dof2node.assign(count, 0);
    m.bkeys.assign(m.nborder, 0);
    for (size_t i = 0; i < m.nelem; i++) {
        if (!m.fkeys[i]) continue;
        m.dof2node[m.dof_id[i]] = (uint32_t)i;
        if (m.fkeys[i] != 255) m.bkeys[m.dof_id[i] - m.nvel] = m.fkeys[i];
    }
}

// ---------------------------------------------------------------------------
// input (femhmg.c:1456-1758)
// ---------------------------------------------------------------------------
bool read_data(Model &m, const char *filename)
{
    FILE *f = fopen(filename, "r");
    if (!f) return false;
    char str[256];
    unsigned count = 0;
    const unsigned num_of_info = 12;
    bool stab_set = false;
    unsigned nx = 0, ny = 0, nz = 0;
    while (count < num_of_info && fscanf(f, "%255s", str) == 1) {
        if (!strcmp(str, "%type_of_analysis")) {
            unsigned a;
            if (fscanf(f, "%u", &a) != 1) break;
            count++;
            if (a > HMG_FLUID) {
                fclose(f);
                printf("ERROR: Provided analysis flag does not match any possible value.\n");
                return false;
            }
            m.analysis = (int)a;
        } else if (!strcmp(str, "%type_of_solver")) {
            unsigned s;
            if (fscanf(f, "%u", &s) != 1) break;
            g_solver = (cudapcgFlag_t)s;
            count++;
        } else if (!strcmp(str, "%voxel_size")) {
            if (fscanf(f, "%lf", &m.elem_size) != 1) break;
            count++;
        } else if (!strcmp(str, "%solver_tolerance")) {
            if (fscanf(f, "%lf", &m.tol) != 1) break;
            count++;
        } else if (!strcmp(str, "%number_of_iterations")) {
            if (fscanf(f, "%u", &m.maxit) != 1) break;
            count++;
        } else if (!strcmp(str, "%image_dimensions")) {
            if (fscanf(f, "%u %u %u", &nx, &ny, &nz) != 3) break;
            count++;
        } else if (!strcmp(str, "%refinement")) {
            if (fscanf(f, "%u", &m.refinement) != 1) break;
            count++;
            if (m.refinement < 1) {
                fclose(f);
                return false;
            }
        } else if (!strcmp(str, "%number_of_materials")) {
            if (fscanf(f, "%u", &m.nmat) != 1) break;
            count++;
            if (m.nmat > 256) {
                fclose(f);
                printf("ERROR: Number of different materials (%d) is greater than max (%d).\n", m.nmat, 255);
                return false;
            }
        } else if (!strcmp(str, "%properties_of_materials")) {
            for (unsigned i = 0; i < m.nmat; i++) {
                unsigned key;
                int ok;
                if (m.analysis == HMG_ELASTIC) ok = fscanf(f, "%u %lf %lf", &key, &m.props[2 * i], &m.props[2 * i + 1]) == 3;
                else if (m.analysis == HMG_THERMAL) ok = fscanf(f, "%u %lf", &key, &m.props[i]) == 2;
                else ok = fscanf(f, "%u", &key) == 1;
                if (!ok) break;
                count++;
                if (key > 255) {
                    fclose(f);
                    printf("ERROR: Material %d has color key (%d) greater than max (%d).\n", i, key, 255);
                    return false;
                }
                m.props_keys[key] = (uint8_t)i;
            }
        } else if (!strcmp(str, "%thermal_expansion")) {
            if (m.analysis == HMG_ELASTIC) {
                for (unsigned i = 0; i < m.nmat; i++)
                    if (fscanf(f, "%lf", &m.alpha[i]) != 1) break;
                m.thermal_expansion = true;
            }
        } else if (!strcmp(str, "%data_type")) {
            if (fscanf(f, "%255s", str) != 1) break;
            m.sdf_f64 = !strcmp(str, "float64");
            count++;
        } else if (!strcmp(str, "%xreduce_scale_factor")) {
            if (fscanf(f, "%lf", &g_xreduce_stab) != 1) break;
            stab_set = true;
            count++;
        }
    }
    fclose(f);
    if (nz == 0) {
        m.dim = HMG_2D;
        printf("ERROR: 2D images are not provided by this library.\n");
        return false;
    }
    m.dim = HMG_3D;
    if (m.nmat > MATKEY_BITSTEP_RANGE_3D + 1 && m.analysis != HMG_FLUID)
        printf("WARNING: Number of specified materials (%d) is greater than what can be represented with "
               "MATKEY_BITSTEP (%d). Either avoid Node-by-Node and DOF-by-DOF, or recompile with greater MATKEY size.\n",
               m.nmat, MATKEY_BITSTEP_3D);
    m.elem_size /= m.refinement;
    m.ex = nx * m.refinement;
    m.ey = ny * m.refinement;
    m.ez = nz * m.refinement;
    if (!stab_set) g_xreduce_stab = m.tol >= 0.0000005 ? m.tol * m.tol : 1e-12;
    return true;
}

bool read_image(Model &m, const uint8_t *img)
{
    const unsigned ref = m.refinement;
    const size_t rows = m.ey / ref, cols = m.ex / ref, slices = m.ez / ref;
    const size_t rows_ref = m.ey, cols_ref = m.ex;
    m.elem_map.assign(m.nelem, 0);
    #pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)slices; k++)
        for (size_t i = 0; i < rows; i++)
            for (size_t j = 0; j < cols; j++) {
                const uint16_t v = m.props_keys[img[(size_t)k * rows * cols + i * cols + j]];
                for (size_t kk = ref * k; kk < ref * (k + 1); kk++)
                    for (size_t ii = ref * i; ii < ref * (i + 1); ii++)
                        for (size_t jj = ref * j; jj < ref * (j + 1); jj++)
                            m.elem_map[ii + jj * rows_ref + kk * rows_ref * cols_ref] = v;
            }
    return true;
}

// strategy of the solve: a scalar density field needs element-by-element
// products (femhmg.c:196-203)
cudapcgFlag_t strategy(const Model &m) { return m.have_sdf ? CUDAPCG_EBE : g_strategy; }

// Scalar density field from a .bin file of float32/float64 values in image
// file order (chfem/src/femhmg/femhmg.c:1760-1835): the element value is the
// file value times the first property of the element's material, quantised
// to 16 bits between the extreme values found (the extremes pair file-order
// values with the element-order material map, as the reference does), and
// replicated over refined elements.  Properties are then normalised: thermal
// runs with one unit material, elastic keeps each material's Poisson ratio
// with E = 1.
bool read_sdf(Model &m, const char *filename)
{
    const unsigned ref = m.refinement;
    const size_t rows = m.ey / ref, cols = m.ex / ref, slices = m.ez / ref;
    const size_t rows_ref = m.ey, cols_ref = m.ex, n0 = rows * cols * slices;
    const size_t sz = m.sdf_f64 ? sizeof(double) : sizeof(float);
    std::vector<unsigned char> raw(sz * m.nelem);
    printf("    Getting density field from .bin...");
    FILE *f = fopen(filename, "rb");
    if (!f) {
        printf(" ERROR: Failed to open %s.\n", filename);
        return false;
    }
    const size_t got = fread(raw.data(), sz, m.nelem, f);
    fclose(f);
    if (got < m.nelem) {
        printf(" ERROR: Failed to read all %e MB from %s.\n", (float)(sz * m.nelem) / 1024.0 / 1024.0, filename);
        return false;
    }
    printf("Done.\n");
    auto val = [&](size_t i) -> double {
        if (m.sdf_f64) {
            double d;
            memcpy(&d, raw.data() + i * sizeof(double), sizeof(double));
            return d;
        }
        float v;
        memcpy(&v, raw.data() + i * sizeof(float), sizeof(float));
        return (double)v;
    };
    const unsigned stride = m.analysis == HMG_ELASTIC ? 2 : 1;
    double Emax = val(0) * m.props[m.elem_map[0] * stride], Emin = Emax;
    for (size_t i = 1; i < n0; i++) {
        const double e = val(i) * m.props[m.elem_map[i] * stride];
        Emax = e > Emax ? e : Emax;
        Emin = e < Emin ? e : Emin;
    }
    if (Emax == Emin) Emin = Emax - 1.0;
    const double dE = Emax - Emin;
    m.dmax = Emax;
    m.dmin = Emin;
    m.density.assign(m.nelem, 0);
    for (size_t k = 0; k < slices; k++)
        for (size_t i = 0; i < rows; i++)
            for (size_t j = 0; j < cols; j++) {
                const double b = val(j + i * cols + k * rows * cols) *
                                 m.props[m.elem_map[(i + j * rows_ref + k * rows_ref * cols_ref) * ref] * stride];
                const uint16_t d = (uint16_t)(65535 * (b - Emin) / dE);
                for (size_t kk = ref * k; kk < ref * (k + 1); kk++)
                    for (size_t ii = ref * i; ii < ref * (i + 1); ii++)
                        for (size_t jj = ref * j; jj < ref * (j + 1); jj++)
                            m.density[ii + jj * rows_ref + kk * rows_ref * cols_ref] = d;
            }
    if (m.analysis == HMG_THERMAL) {
        m.nmat = 1;
        m.props[0] = 1.0;
        std::fill(m.elem_map.begin(), m.elem_map.end(), (uint16_t)0);
    } else {
        for (unsigned j = 0; j < m.nmat; j++) m.props[2 * j] = 1.0;
    }
    m.have_sdf = true;
    return true;
}

bool read_raw(Model &m, const char *filename)
{
    const size_t n = (size_t)(m.ex / m.refinement) * (m.ey / m.refinement) * (m.ez / m.refinement);
    FILE *f = fopen(filename, "rb");
    if (!f) {
        printf(" ERROR: Failed to open file.\n");
        return false;
    }
    std::vector<uint8_t> buf(n);
    const size_t got = fread(buf.data(), 1, n, f);
    fclose(f);
    if (got < n) {
        printf(" ERROR: Reached end of file before expected.\n");
        return false;
    }
    return read_image(m, buf.data());
}

// model -> solver model (femhmg.c:54-85)
cudapcgFlag_t model_ctor(cudapcgModel_t **ptr, const void *data)
{
    if (!data) return CUDAPCG_FALSE;
    const Model *m = (const Model *)data;
    if (m->nmat == 0 || m->nelem == 0) return CUDAPCG_FALSE;
    cudapcgModel_t *p = *ptr;
    p->name = (char *)malloc(m->nf_noext.size() + 1);
    strcpy(p->name, m->nf_noext.c_str());
    p->parStrategy_flag = strategy(*m);
    p->poremap_flag = m->poremap;
    p->parametric_density_field_flag = m->have_sdf ? CUDAPCG_TRUE : CUDAPCG_FALSE;
    p->nrows = m->ey;
    p->ncols = m->ex;
    p->nlayers = m->ez;
    p->nelem = (unsigned)m->nelem;
    p->nvars = (unsigned)m->ndof;
    p->nvarspernode = m->dim;
    p->nkeys = m->analysis == HMG_FLUID ? 1 : m->nmat;
    p->localmtxdim = m->lclMtx_dim;
    p->nporenodes = m->nvel;
    p->nbordernodes = m->nborder;
    p->nhmgvars = m->analysis == HMG_FLUID ? m->nvel * m->dim : (unsigned)m->ndof;
    return CUDAPCG_TRUE;
}

void appendf(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void appendf(const char *fmt, ...)
{
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    printf("%s", buf);
    if (M && M->report_on) M->report.push_back(buf);
}

// load case table: seam axis, seam local nodes, jumped component
// (femhmg_thermal_3D.c:144-234, femhmg_elastic_3D.c:152-462)
LoadCase load_case(const Model &m, unsigned dir)
{
    static const int AX[6] = {0, 1, 2, 1, 2, 1};
    static const int SN[6][4] = {{1, 2, 5, 6}, {2, 3, 6, 7}, {0, 1, 2, 3}, {2, 3, 6, 7}, {0, 1, 2, 3}, {2, 3, 6, 7}};
    static const int CMP[6] = {0, 1, 2, 2, 0, 0};
    LoadCase lc;
    lc.axis = AX[dir];
    for (int t = 0; t < 4; t++) lc.S[t] = SN[dir][t];
    lc.comp = m.analysis == HMG_ELASTIC ? CMP[dir] : 0;
    lc.L = lc.axis == 0 ? m.ex : (lc.axis == 1 ? m.ey : m.ez);
    if (m.analysis == HMG_FLUID) {
        double f = 0;
        for (int k = 0; k < 8; k++) f += m.CB[72 + k];
        lc.L = f;
    }
    return lc;
}

struct DeviceAux {
    double *CB = nullptr, *seam = nullptr, *K = nullptr;
    double *F = nullptr;      // corner field matrices (field export)
    double *Bc = nullptr;     // centre strain matrix per material (thermal expansion)
    double *da = nullptr;     // per material (D alpha)_xx (thermal expansion)
    double *fout = nullptr;   // nodal field output
};

bool upload_aux(const Model &m, DeviceAux &a)
{
    if (m.analysis == HMG_FLUID) return true;
    const std::vector<double> K = dense_K(m);
    if (hipMalloc(&a.CB, m.CB.size() * sizeof(double)) != hipSuccess) return false;
    if (hipMalloc(&a.seam, 6 * m.nmat * sizeof(double)) != hipSuccess) return false;
    if (hipMalloc(&a.K, K.size() * sizeof(double)) != hipSuccess) return false;
    hipMemcpy(a.CB, m.CB.data(), m.CB.size() * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(a.K, K.data(), K.size() * sizeof(double), hipMemcpyHostToDevice);
    return true;
}

void free_aux(DeviceAux &a)
{
    if (a.F) hipFree(a.F);
    if (a.Bc) hipFree(a.Bc);
    if (a.da) hipFree(a.da);
    if (a.fout) hipFree(a.fout);
    if (a.CB) hipFree(a.CB);
    if (a.seam) hipFree(a.seam);
    if (a.K) hipFree(a.K);
    a = DeviceAux();
}

// seam term of the C column per material and row (6 * nmat entries): the sum
// of CB over the load case's seam columns, times the jump
void seam_table(const Model &m, const LoadCase &lc, double *seam)
{
    const int nd = m.analysis == HMG_ELASTIC ? 3 : 1, ke = 8 * nd, nr = nd == 3 ? 6 : 3;
    for (size_t i = 0; i < 6 * (size_t)m.nmat; i++) seam[i] = 0.0;
    for (unsigned t = 0; t < m.nmat; t++)
        for (int row = 0; row < nr; row++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += m.CB[(size_t)t * nr * ke + row * ke + nd * lc.S[k] + lc.comp];
            seam[t * 6 + row] = s * lc.L;
        }
}

bool set_seam(const Model &m, const DeviceAux &a, const LoadCase &lc)
{
    if (m.analysis == HMG_FLUID) return true;
    std::vector<double> seam(6 * m.nmat, 0.0);
    seam_table(m, lc, seam.data());
    return hipMemcpy(a.seam, seam.data(), seam.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
}

// RHS for direction dir into the solver's RHS buffer
bool device_rhs(const Model &m, const DeviceAux &a, unsigned dir)
{
    const LoadCase lc = load_case(m, dir);
    DeviceModel *dm = active_model();
    launch_rhs(dm->op64, a.K, lc, (int)dir, rhs_buffer(), active_stream());
    return hipStreamSynchronize(active_stream()) == hipSuccess;
}

// column dir of C from the solution in the solver's x buffer
bool device_updateC(Model &m, const DeviceAux &a, unsigned dir, const double *x)
{
    const LoadCase lc = load_case(m, dir);
    DeviceModel *dm = active_model();
    Reduce red = make_reduce(FIN_STORE, S_TMP0);
    double xr[4];
    if (m.analysis == HMG_FLUID && reduced_solution(xr) >= 3) {
        // CG2 / MINRES2: the solver reduced the velocity on the fly and
        // returns the column itself (femhmg.c:1087-1093)
        for (int row = 0; row < 3; row++) m.C[row * 3 + dir] = xr[row];
        return true;
    }
    if (m.analysis == HMG_FLUID) {
        launch_fluid_velocity_sum<double>(dm->op64, x, &red, active_stream());
        if (!sync_scalars()) return false;
        const double *S = scalars_host();
        const double mul = m.elem_size * m.elem_size / (double)m.nelem;
        for (int row = 0; row < 3; row++) m.C[row * 3 + dir] = S[S_TMP0 + row] * mul;
        return true;
    }
    if (!set_seam(m, a, lc)) return false;
    launch_updateC(dm->op64, a.CB, a.seam, lc, x, &red, active_stream());
    if (!sync_scalars()) return false;
    const double *S = scalars_host();
    const int nr = m.analysis == HMG_ELASTIC ? 6 : 3;
    for (int row = 0; row < nr; row++) m.C[row * nr + dir] = S[S_TMP0 + row] / (double)m.nelem;
    return true;
}

// ---------------------------------------------------------------------------
// field export (femhmg_thermal_3D.c:694-1105, femhmg_elastic_3D.c:1800-2645,
// femhmg_fluid_3D.c:312-360)
// ---------------------------------------------------------------------------
// grad N_j at corner k of the unit voxel (element frame, z = 1 - layer offset):
// only corners sharing the other two coordinates with j contribute
void corner_grad(int k, int j, double g[3])
{
    int xk, yk, zk, xj, yj, zj;
    frame(k, false, xk, yk, zk);
    frame(j, false, xj, yj, zj);
    g[0] = (yj == yk && zj == zk) ? (xj ? 1.0 : -1.0) : 0.0;
    g[1] = (xj == xk && zj == zk) ? (yj ? 1.0 : -1.0) : 0.0;
    g[2] = (xj == xk && yj == yk) ? (zj ? 1.0 : -1.0) : 0.0;
}

// strain (Voigt, engineering shear) of displacement component b of node j
// for gradient g of N_j
void strain_of(const double g[3], int b, double eps[6])
{
    for (int r = 0; r < 6; r++) eps[r] = 0;
    eps[b] = g[b];
    if (b == 0) { eps[4] = g[2]; eps[5] = g[1]; }
    if (b == 1) { eps[3] = g[2]; eps[5] = g[0]; }
    if (b == 2) { eps[3] = g[1]; eps[4] = g[0]; }
}

// per material, per corner k: thermal k grad N (3x8), elastic D B (6x24)
std::vector<double> corner_field_matrices(const Model &m)
{
    const bool el = m.analysis == HMG_ELASTIC;
    const int nd = el ? 3 : 1, ke = 8 * nd, nf = el ? 6 : 3, sz = nf * ke;
    std::vector<double> F((size_t)m.nmat * 8 * sz, 0.0);
    for (unsigned t = 0; t < m.nmat; t++)
        for (int k = 0; k < 8; k++) {
            double *Fk = &F[((size_t)t * 8 + k) * sz];
            for (int j = 0; j < 8; j++) {
                double g[3];
                corner_grad(k, j, g);
                if (!el) {
                    for (int a = 0; a < 3; a++) Fk[a * ke + j] = m.props[t] * g[a];
                    continue;
                }
                const double E = m.props[2 * t], nu = m.props[2 * t + 1];
                const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
                for (int b = 0; b < 3; b++) {
                    double eps[6];
                    strain_of(g, b, eps);
                    const double tr = eps[0] + eps[1] + eps[2];
                    for (int r = 0; r < 3; r++) Fk[r * ke + 3 * j + b] = lam * tr + 2 * mu * eps[r];
                    for (int r = 3; r < 6; r++) Fk[r * ke + 3 * j + b] = mu * eps[r];
                }
            }
        }
    return F;
}

bool write_bin(const std::string &name, const double *v, size_t n)
{
    FILE *f = fopen(name.c_str(), "wb");
    if (!f) {
        printf("WARNING: could not write %s\n", name.c_str());
        return false;
    }
    const bool ok = fwrite(v, sizeof(double), n, f) == n;
    fclose(f);
    return ok;
}

// writes <nf>_{temperature,flux|displacement,stress|velocity,pressure}_<dir>.bin
// for the solution in the solver's x buffer
bool export_fields(const Model &m, DeviceAux &a, unsigned dir)
{
    const double t0 = omp_get_wtime();
    printf("    Exporting simulation results...");
    std::vector<double> x(m.ndof);
    if (hipMemcpy(x.data(), solution_buffer(), m.ndof * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    const std::string base = m.nf_noext + "_";
    const std::string sfx = "_" + std::to_string(dir) + ".bin";
    bool ok;
    if (m.analysis == HMG_FLUID) {
        std::vector<double> vel(3 * m.nelem, 0.0), pres(m.nelem, 0.0);
#pragma omp parallel for schedule(static)
        for (long n = 0; n < (long)m.nelem; n++) {
            const uint32_t id = m.dof_id[n];
            if (m.fkeys[n] == 255)
                for (int c = 0; c < 3; c++) vel[3 * n + c] = x[3 * (size_t)id + c];
            if (m.fkeys[n]) pres[n] = x[3 * (size_t)m.nvel + id];
        }
        ok = write_bin(base + "velocity" + sfx, vel.data(), vel.size()) &&
             write_bin(base + "pressure" + sfx, pres.data(), pres.size());
    } else {
        const bool el = m.analysis == HMG_ELASTIC;
        const size_t nf = el ? 6 : 3;
        if (!a.F) {
            const std::vector<double> F = corner_field_matrices(m);
            if (hipMalloc(&a.F, F.size() * sizeof(double)) != hipSuccess ||
                hipMemcpy(a.F, F.data(), F.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
                return false;
        }
        if (!a.fout && hipMalloc(&a.fout, nf * m.nelem * sizeof(double)) != hipSuccess) return false;
        DeviceModel *dm = active_model();
        launch_fields(dm->op64, a.F, load_case(m, dir), 1, solution_buffer(), a.fout, active_stream());
        std::vector<double> fld(nf * m.nelem);
        if (hipStreamSynchronize(active_stream()) != hipSuccess ||
            hipMemcpy(fld.data(), a.fout, fld.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            return false;
        ok = write_bin(base + (el ? "displacement" : "temperature") + sfx, x.data(), x.size()) &&
             write_bin(base + (el ? "stress" : "flux") + sfx, fld.data(), fld.size());
    }
    printf("Done.(%.2e s)\n", omp_get_wtime() - t0);
    return ok;
}

// ---------------------------------------------------------------------------
// thermal expansion (femhmg.c:1121-1165, femhmg_elastic_3D.c:464-508,
// 863-1073): solve K u = sum_e B_c^T D_e a_e (no seam jump) and
// alpha* = C^-1 beta with beta = <D_e (a_e - B_c u_e)> + C <B_c u_e>.
// ---------------------------------------------------------------------------
bool thermal_expansion_pass(Model &m, DeviceAux &a)
{
    appendf("    ------------------------------------------------------\n    Thermal expansion analysis:\n");
    const size_t nm = m.nmat;
    std::vector<double> da(nm), Bc(144 * nm);
    for (unsigned t = 0; t < nm; t++) {
        const double E = m.props[2 * t], nu = m.props[2 * t + 1];
        const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
        da[t] = (3 * lam + 2 * mu) * m.alpha[t];  // (D a)_xx for a = alpha (1,1,1,0,0,0)
    }
    // centre strain matrix (6x24), the same for every material
    for (unsigned t = 0; t < nm; t++)
        for (int k = 0; k < 8; k++) {
            int x, y, z;
            frame(k, false, x, y, z);
            const double g[3] = {(x ? 0.25 : -0.25), (y ? 0.25 : -0.25), (z ? 0.25 : -0.25)};
            for (int b = 0; b < 3; b++) {
                double eps[6];
                strain_of(g, b, eps);
                for (int r = 0; r < 6; r++) Bc[144 * t + r * 24 + 3 * k + b] = eps[r];
            }
        }
    if ((!a.da && hipMalloc(&a.da, nm * sizeof(double)) != hipSuccess) ||
        (!a.Bc && hipMalloc(&a.Bc, Bc.size() * sizeof(double)) != hipSuccess))
        return false;
    hipMemcpy(a.da, da.data(), nm * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(a.Bc, Bc.data(), Bc.size() * sizeof(double), hipMemcpyHostToDevice);
    hipMemset(a.seam, 0, 6 * nm * sizeof(double));
    DeviceModel *dm = active_model();
    launch_rhs_thermal_expansion(dm->op64, a.da, rhs_buffer(), active_stream());
    if (hipStreamSynchronize(active_stream()) != hipSuccess) return false;
    cudapcgSetX0(nullptr, CUDAPCG_FALSE);
    const int rc = solve_device(nullptr);
    chfemSolveStats_t st;
    cudapcgGetSolveStats(&st);
    m.stats.push_back(st);
    appendf("    | Solver ran %u iterations (%u sweeps), residual %.3e, %.3e s\n", st.iterations, st.sweeps,
            st.residual, st.solve_seconds);
    if (rc < 0) return false;
    // <D_e B_c u_e> and <B_c u_e>: the C-column kernel with a zero seam term
    const LoadCase lc0 = load_case(m, 0);
    double sum_s[6], sum_e[6];
    for (int pass = 0; pass < 2; pass++) {
        Reduce red = make_reduce(FIN_STORE, S_TMP0);
        launch_updateC(dm->op64, pass ? a.Bc : a.CB, a.seam, lc0, solution_buffer(), &red, active_stream());
        if (!sync_scalars()) return false;
        const double *S = scalars_host();
        for (int r = 0; r < 6; r++) (pass ? sum_e : sum_s)[r] = S[S_TMP0 + r];
    }
    std::vector<size_t> count(nm, 0);
    for (size_t e = 0; e < m.nelem; e++) count[m.elem_map[e]]++;
    double da_sum = 0;
    for (unsigned t = 0; t < nm; t++) da_sum += (double)count[t] * da[t];
    double stress[6], strain[6], beta[6], Mx[36];
    for (int r = 0; r < 6; r++) {
        stress[r] = (sum_s[r] - (r < 3 ? da_sum : 0.0)) / (double)m.nelem;
        strain[r] = sum_e[r] / (double)m.nelem;
    }
    for (int i = 0; i < 6; i++) {
        beta[i] = -stress[i];
        for (int j = 0; j < 6; j++) beta[i] += m.C[6 * i + j] * strain[j];
    }
    // alpha* = C^-1 beta, Gaussian elimination without pivoting (C is SPD)
    memcpy(Mx, m.C, sizeof(Mx));
    for (int i = 0; i < 5; i++)
        for (int j = i + 1; j < 6; j++) {
            const double f = Mx[6 * j + i] / Mx[7 * i];
            for (int k = i + 1; k < 6; k++) Mx[6 * j + k] -= Mx[6 * i + k] * f;
            beta[j] -= beta[i] * f;
            Mx[6 * j + i] = 0;
        }
    for (int i = 5; i >= 0; i--) {
        double v = beta[i];
        for (int j = 5; j > i; j--) v -= Mx[6 * i + j] * m.texp[j];
        m.texp[i] = v / Mx[7 * i];
    }
    m.texp_ok = true;
    return true;
}

void directions(const Model &m, std::vector<unsigned> &dirs)
{
    dirs.clear();
    if (m.hmg_flag_set) dirs.push_back(m.hmg_flag);
    else if (m.analysis == HMG_ELASTIC) for (unsigned d = 0; d < 6; d++) dirs.push_back(d);
    else for (unsigned d = 0; d < 3; d++) dirs.push_back(d);
}

// set up the solver for model m (femhmg.c:914-1002)
bool solver_setup(Model &m, DeviceAux &aux)
{
    if (!cudapcgInit(m.pcg_flag, m.analysis == HMG_FLUID ? CUDAPCG_NBN : strategy(m))) return false;
    cudapcgSetXReduceMode(g_xreduce);
    cudapcgSetXReduceScale(m.elem_size * m.elem_size / (double)m.nelem);
    cudapcgSetReduceStabFactor(g_xreduce_stab);
    cudapcgSetHeaderString((char *)"    | ");
    cudapcgSetModelConstructorFcn(model_ctor);
    if (!cudapcgBuildModel(&m)) { cudapcgEnd(); return false; }
    if (!cudapcgSetSolver(g_solver) || !cudapcgSetPreconditioner(g_precond) || !cudapcgSetResNorm(g_stopcrit)) {
        cudapcgEnd();
        return false;
    }
    cudapcgSetPrecision(g_precision);
    cudapcgSetNumTol(m.tol);
    cudapcgSetMaxIterations(m.maxit);
    if (m.analysis != HMG_FLUID) {
        if (strategy(m) == CUDAPCG_NBN) cudapcgSetImage(m.keys.data());
        else cudapcgSetImage(m.elem_map.data());
        if (m.have_sdf && !cudapcgSetParametricDensityField(m.density.data(), m.dmin, m.dmax)) {
            cudapcgEnd();
            return false;
        }
    } else {
        cudapcgSetPeriodic2DOFMap(m.dof_id.data());
        cudapcgSetPoreMap(m.bkeys.data());
        cudapcgSetDOF2PeriodicMap(m.dof2node.data());
    }
    if (!cudapcgSetLclMtxs(m.Mtxs.data())) { cudapcgEnd(); return false; }
    if (!cudapcgAllocateArrays()) { cudapcgEnd(); return false; }
    if (!upload_aux(m, aux)) { cudapcgEnd(); return false; }
    return true;
}

bool init_sizes(Model &m)
{
    if (m.ex == 0 || m.ey == 0 || m.ez == 0 || m.nmat == 0) {
        printf("ERROR: Failed to initialize model due to bad input parameters.\nrows:%d\ncols:%d\nslices:%d\nmaterials:%d\n",
               m.ey, m.ex, m.ez, m.nmat);
        return false;
    }
    m.nelem = (size_t)m.ex * m.ey * m.ez;
    if (m.nelem >= 0xffffffffull) {
        printf("ERROR: image too large for 32-bit node indices.\n");
        return false;
    }
    if (m.analysis == HMG_THERMAL) {
        m.C_dim = 9; m.lclMtx_dim = 40; m.lclCB_dim = 0; m.ndof = m.nelem; m.pcg_flag = CUDAPCG_THERMAL_3D;
    } else if (m.analysis == HMG_ELASTIC) {
        m.C_dim = 36; m.lclMtx_dim = 576; m.lclCB_dim = 144; m.ndof = 3 * m.nelem; m.pcg_flag = CUDAPCG_ELASTIC_3D;
    } else {
        m.C_dim = 9; m.lclMtx_dim = 1024; m.lclCB_dim = 80; m.ndof = m.nelem; m.pcg_flag = CUDAPCG_FLUID_3D;
    }
    return true;
}

}  // namespace

extern "C" {

logical hmgInit(char *data_filename, char *elem_filename, char *sdf_filename, uint8_t *npdata)
{
    g_time_total = omp_get_wtime();
    printf("Initializing femhmg...\n");
    if (M) hmgEnd();
    if (!data_filename) return HMG_FALSE;
    M = new Model();
    Model &m = *M;
    memset(m.props, 0, sizeof(m.props));
    memset(m.alpha, 0, sizeof(m.alpha));
    memset(m.props_keys, 0, sizeof(m.props_keys));
    memset(m.C, 0, sizeof(m.C));
    m.nf = data_filename;
    m.nf_noext = m.nf.size() >= 3 ? m.nf.substr(0, m.nf.size() - 3) : m.nf;
    if (elem_filename) m.image_file = elem_filename;
    m.poremap = g_poremap_default;
    if (!read_data(m, data_filename)) { delete M; M = nullptr; return HMG_FALSE; }
    if (sdf_filename && m.analysis == HMG_FLUID) {
        printf("WARNING: Scalar density field input is not supported for permeability analysis.\n");
        sdf_filename = nullptr;
    }
    if (!init_sizes(m)) { delete M; M = nullptr; return HMG_FALSE; }
    bool ok;
    if (!elem_filename && npdata) ok = read_image(m, npdata);
    else if (elem_filename && strlen(elem_filename) > 4 && !strcmp(elem_filename + strlen(elem_filename) - 4, ".raw"))
        ok = read_raw(m, elem_filename);
    else {
        printf("ERROR: file %s is not in a valid format.\n", elem_filename ? elem_filename : "(null)");
        ok = false;
    }
    if (ok && sdf_filename) ok = read_sdf(m, sdf_filename);
    if (!ok) { delete M; M = nullptr; return HMG_FALSE; }
    if (m.analysis == HMG_FLUID) build_fluid_maps(m);
    else build_keys(m);
    assemble_local(m);
    printf("Done.(%.2e s)\n", omp_get_wtime() - g_time_total);
    return HMG_TRUE;
}

logical hmgEnd(void)
{
    if (!M) return HMG_FALSE;
    hmgRelease();
    delete M;
    M = nullptr;
    printf("Elapsed time (total): %.2e s\n", omp_get_wtime() - g_time_total);
    return HMG_TRUE;
}

void hmgSetXReduceFlag(cudapcgFlag_t flag) { g_xreduce = flag <= CUDAPCG_XREDUCE_FULL ? flag : CUDAPCG_XREDUCE_FULL; }
void hmgSetParallelStrategyFlag(cudapcgFlag_t flag) { g_strategy = flag ? CUDAPCG_EBE : CUDAPCG_NBN; }
void hmgSetSolverFlag(cudapcgFlag_t flag) { g_solver = flag; }
void hmgSetPreConditionerFlag(cudapcgFlag_t flag) { g_precond = flag > 0 ? CUDAPCG_TRUE : CUDAPCG_FALSE; }
void hmgSetStoppingCriteria(cudapcgFlag_t flag) { g_stopcrit = flag; }
void hmgSetPrecisionFlag(cudapcgFlag_t flag) { g_precision = flag ? CHFEM_PREC_MIXED : CHFEM_PREC_FP64; }

logical hmgSetPoreMappingStrategy(cudapcgFlag_t flag)
{
    if (!M) {
        g_poremap_default = flag;
        return HMG_TRUE;
    }
    if (flag == M->poremap || flag > CUDAPCG_POREMAP_NUM) return HMG_TRUE;
    if (M->analysis != HMG_FLUID) {
        M->poremap = flag;
        return HMG_TRUE;
    }
    printf("ERROR: image pore mapping is not provided by this library; use DOF-number mapping.\n");
    return HMG_FALSE;
}

void hmgSetHomogenizationFlag(hmgFlag_t flag)
{
    if (!M) return;
    if (flag > HOMOGENIZE_XY) { M->hmg_flag_set = false; return; }
    if ((M->analysis == HMG_THERMAL || M->analysis == HMG_FLUID) && flag > HOMOGENIZE_Z) {
        M->hmg_flag_set = false;
        return;
    }
    M->hmg_flag_set = true;
    M->hmg_flag = flag;
}

void hmgExportX(logical flag) { if (M) M->export_x = flag > 0; }
void hmgImportX(char *file)
{
    if (!M || !file) return;
    M->x0_file = file;
    M->using_x0 = true;
}

// Recursive coarse-mesh initial guesses (femhmg.c:561-846): every level halves
// the element map (taking every second voxel, femhmg.c:1844-1916), is solved
// with the fine model's settings, and its solution is interpolated onto the
// next finer level through cudapcgSetX0(x0, mustInterpolate=TRUE).
void hmgFindInitialGuesses(unsigned int nlevels)
{
    if (!nlevels || !M) return;
    if (M->have_sdf) {
        printf("WARNING: Initial guess estimation is not available with a scalar density field input.\n");
        return;
    }
    if (M->analysis == HMG_FLUID) {
        printf("WARNING: Initial guess estimation is not yet available for permeability analysis.\n");
        printf("         Null initial guess will be considered (x0 = [0]).\n");
        return;
    }
    if (g_solver > CUDAPCG_MINRES3_SOLVER)
        g_solver = g_solver == CUDAPCG_CG2_SOLVER ? CUDAPCG_CG3_SOLVER : CUDAPCG_MINRES3_SOLVER;
    std::vector<Model *> lv;
    lv.push_back(M);
    for (unsigned i = 0; i < nlevels; i++) {
        const Model &f = *lv.back();
        Model *c = new Model(f);
        c->ex = f.ex / 2; c->ey = f.ey / 2; c->ez = f.ez / 2;
        c->report_on = false;
        c->x0.clear();
        c->using_x0 = false;
        c->x0_file.clear();
        c->export_x = false;
        if (c->ex == 0 || c->ey == 0 || c->ez == 0) { delete c; break; }
        c->nelem = (size_t)c->ex * c->ey * c->ez;
        c->ndof = c->analysis == HMG_ELASTIC ? 3 * c->nelem : c->nelem;
        c->elem_map.assign(c->nelem, 0);
        for (unsigned r = 0; r < c->ey; r++)
            for (unsigned col = 0; col < c->ex; col++)
                for (unsigned l = 0; l < c->ez; l++)
                    c->elem_map[r + (size_t)col * c->ey + (size_t)l * c->ex * c->ey] =
                        f.elem_map[2 * r + (size_t)2 * col * f.ey + (size_t)2 * l * f.ex * f.ey];
        build_keys(*c);
        lv.push_back(c);
    }
    std::vector<unsigned> dirs;
    directions(*M, dirs);
    for (size_t i = lv.size() - 1; i > 0; i--) {
        Model &c = *lv[i];
        Model &finer = *lv[i - 1];
        printf("Computing initial guesses in coarse mesh %zu...\n", i);
        DeviceAux aux;
        if (!solver_setup(c, aux)) continue;
        finer.x0.assign(6, std::vector<double>());
        for (unsigned d : dirs) {
            if (!c.x0.empty() && !c.x0[d].empty()) cudapcgSetX0(c.x0[d].data(), CUDAPCG_TRUE);
            else cudapcgSetX0(nullptr, CUDAPCG_FALSE);
            device_rhs(c, aux, d);
            if (solve_device(nullptr) == 1) {
                finer.x0[d].resize(c.ndof);
                hipMemcpy(finer.x0[d].data(), solution_buffer(), c.ndof * sizeof(double), hipMemcpyDeviceToHost);
            }
        }
        free_aux(aux);
        cudapcgEnd();
        if (i > 0) finer.using_x0 = true;
    }
    for (size_t i = 1; i < lv.size(); i++) delete lv[i];
}

void hmgSetSolverTolerance(double tol) { if (M) M->tol = tol; }
void hmgSetMaxIterationsOverride(unsigned int maxit) { if (M) M->maxit = maxit; }

static DeviceAux g_aux;
static bool g_prepared = false;
static double g_op_ms = 0;
static unsigned g_op_count = 0;

logical hmgPrepare(void)
{
    if (!M) return HMG_FALSE;
    if (g_prepared) hmgRelease();
    if (M->analysis != HMG_FLUID && g_solver > CUDAPCG_MINRES3_SOLVER) {
        printf("WARNING: CG2 and MINRES2 are only available for FLUID.\n");
        g_solver = g_solver == CUDAPCG_CG2_SOLVER ? CUDAPCG_CG3_SOLVER : CUDAPCG_MINRES3_SOLVER;
    }
    if (!solver_setup(*M, g_aux)) return HMG_FALSE;
    g_prepared = true;
    // set-up for the allocation-free entry point: operator descriptors
    // built, and the scratch array's device address looked up on this device
    if (!active_model() || !noalloc_scratch()) {
        hmgRelease();
        return HMG_FALSE;
    }
    return HMG_TRUE;
}

void hmgRelease(void)
{
    if (!g_prepared) return;
    free_aux(g_aux);
    cudapcgFreeArrays();
    cudapcgSetX0(nullptr, CUDAPCG_FALSE);
    cudapcgEnd();
    g_prepared = false;
}

logical hmgSolveResident(var *Cout)
{
    if (!M || !g_prepared) return HMG_FALSE;
    Model &m = *M;
    for (unsigned i = 0; i < m.C_dim; i++) m.C[i] = 0.0;
    m.stats.clear();
    cudapcgSetKernelTiming(CUDAPCG_TRUE);
    static const char sim_str[6][11] = {"X", "Y", "Z", "YZ (shear)", "XZ (shear)", "XY (shear)"};
    std::vector<unsigned> dirs;
    directions(m, dirs);
    std::vector<double> xfile;
    bool ok = true;
    // learned initial guess: one prediction for all load cases
    const float *gy = nullptr;
    int gC = 0;
    double *x0dev = nullptr;
    // the guesses of every load case, formed in one pass over the network's
    // output (a vector per load case; if they do not fit, one vector filled
    // per load case)
    double *x0all[6] = {};
    const int gnd = m.analysis == HMG_ELASTIC ? 3 : 1;
    if (guess_active && !m.using_x0 && (m.analysis == HMG_THERMAL || m.analysis == HMG_ELASTIC)) {
        gy = guess_predict_device(guess_active, m.elem_map.data(), m.props, m.nmat, m.ex, m.ey, m.ez, &gC);
        if (gy && hipStreamSynchronize(guess_stream(guess_active)) != hipSuccess) gy = nullptr;
        if (gy) {
            bool all = true;
            for (unsigned d : dirs)
                if (d < 6 && hipMalloc(&x0all[d], m.ndof * sizeof(double)) != hipSuccess) {
                    (void)hipGetLastError();
                    all = false;
                    break;
                }
            if (all) {
                nn::guess_to_x0_all(gy, gC, gnd, nn::Grid3{(int)m.ey, (int)m.ex, (int)m.ez}, 1.0f, x0all,
                                    active_stream());
            } else {
                for (double *&p : x0all) {
                    if (p) (void)hipFree(p);
                    p = nullptr;
                }
                if (hipMalloc(&x0dev, m.ndof * sizeof(double)) != hipSuccess) gy = nullptr;
            }
        }
        if (!gy) appendf("    | learned initial guess unavailable for this grid: zero guess\n");
    }
    for (unsigned d : dirs) {
        appendf("    ------------------------------------------------------\n    Analysis on %s:\n", sim_str[d]);
        m.hmg_flag = d;
        device_rhs(m, g_aux, d);
        if (m.using_x0) {
            if (m.x0_file.empty()) {
                if (!m.x0.empty() && !m.x0[d].empty()) cudapcgSetX0(m.x0[d].data(), CUDAPCG_TRUE);
                else cudapcgSetX0(nullptr, CUDAPCG_FALSE);
            } else {
                char fb[2048];
                snprintf(fb, sizeof(fb), "%s_%d.bin", m.x0_file.c_str(), d);
                FILE *f = fopen(fb, "rb");
                if (!f) {
                    printf("    | Reading initial guess from %s...Failed!...\n", fb);
                    cudapcgSetX0(nullptr, CUDAPCG_FALSE);
                } else {
                    xfile.assign(m.ndof, 0.0);
                    if (fread(xfile.data(), sizeof(double), m.ndof, f) < m.ndof)
                        printf("WARNING: Failed to read all %zu items from %s...", m.ndof, fb);
                    fclose(f);
                    cudapcgSetX0(xfile.data(), CUDAPCG_FALSE);
                }
            }
        }
        int rc;
        if (gy && d < 6 && x0all[d]) {
            rc = solve_device(x0all[d]);
        } else if (gy && x0dev) {
            nn::guess_to_x0(gy, gC, (int)d, gnd, nn::Grid3{(int)m.ey, (int)m.ex, (int)m.ez}, 1.0f, x0dev,
                            active_stream());
            rc = solve_device(x0dev);
        } else {
            rc = solve_device(nullptr);
        }
        chfemSolveStats_t st;
        cudapcgGetSolveStats(&st);
        m.stats.push_back(st);
        appendf("    | Solver ran %u iterations (%u sweeps), residual %.3e, %.3e s\n", st.iterations, st.sweeps,
                st.residual, st.solve_seconds);
        if (rc < 0) {
            ok = false;
            break;
        }
        if (m.export_x && g_solver >= CUDAPCG_CG2_SOLVER) {
            printf("    WARNING: CG2 and MINRES2 keep no solution vector to export.\n");
        } else if (m.export_x) {
            std::vector<double> xh(m.ndof);
            hipMemcpy(xh.data(), solution_buffer(), m.ndof * sizeof(double), hipMemcpyDeviceToHost);
            char fb[2048];
            snprintf(fb, sizeof(fb), "%s_X_%d.bin", m.nf_noext.c_str(), d);
            FILE *f = fopen(fb, "wb");
            if (f) {
                fwrite(xh.data(), sizeof(double), m.ndof, f);
                fclose(f);
            }
        }
        if (!device_updateC(m, g_aux, d, solution_buffer())) {
            ok = false;
            break;
        }
        if (g_solver >= CUDAPCG_CG2_SOLVER && m.analysis == HMG_FLUID) {
            printf("    Effective fields:\n");
            for (unsigned jj = 0; jj < 3; jj++) printf("        %.8e\n", m.C[jj * 3 + d]);
        } else if (m.save_fields && !export_fields(m, g_aux, d)) {
            ok = false;
            break;
        }
    }
    if (x0dev) (void)hipFree(x0dev);
    for (double *p : x0all)
        if (p) (void)hipFree(p);
    m.texp_ok = false;
    if (ok && m.thermal_expansion) {
        if (m.analysis != HMG_ELASTIC)
            printf("    WARNING: Thermal expansion analysis only available for ELASTIC.\n");
        else if (m.have_sdf)
            printf("    WARNING: Thermal expansion analysis is not available with a scalar density field.\n");
        else if (dirs.size() < 6)
            printf("    WARNING: Thermal expansion analysis needs the six load cases (direction 'all').\n");
        else if (!thermal_expansion_pass(m, g_aux))
            ok = false;
    }
    cudapcgGetKernelTiming(&g_op_ms, &g_op_count);
    appendf("    ------------------------------------------------------\n");
    if (Cout) memcpy(Cout, m.C, m.C_dim * sizeof(double));
    return ok ? HMG_TRUE : HMG_FALSE;
}

logical hmgSolveLoadCase(unsigned int dir, const var *x0, var *x, var *Ccol)
{
    if (!M || !g_prepared || dir >= (M->analysis == HMG_ELASTIC ? 6u : 3u)) return HMG_FALSE;
    Model &m = *M;
    m.hmg_flag = dir;
    device_rhs(m, g_aux, dir);
    int rc;
    if (x0) {
        double *xd = nullptr;
        if (hipMalloc(&xd, m.ndof * sizeof(double)) != hipSuccess) return HMG_FALSE;
        if (hipMemcpy(xd, x0, m.ndof * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(xd);
            return HMG_FALSE;
        }
        rc = solve_device(xd);
        (void)hipFree(xd);
    } else {
        cudapcgSetX0(nullptr, CUDAPCG_FALSE);
        rc = solve_device(nullptr);
    }
    if (rc < 0) return HMG_FALSE;
    if (x && g_solver >= CUDAPCG_CG2_SOLVER) {
        fprintf(stderr, "chfem: CG2 and MINRES2 keep no solution vector\n");
        return HMG_FALSE;
    }
    if (x && hipMemcpy(x, solution_buffer(), m.ndof * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return HMG_FALSE;
    if (Ccol) {
        for (unsigned i = 0; i < m.C_dim; i++) m.C[i] = 0.0;
        if (!device_updateC(m, g_aux, dir, solution_buffer())) return HMG_FALSE;
        const unsigned n = m.analysis == HMG_ELASTIC ? 6 : 3;
        for (unsigned i = 0; i < n; i++) Ccol[i] = m.C[i * n + dir];
    }
    return rc == 1 ? HMG_TRUE : HMG_FALSE;
}

cudapcgFlag_t hmgGetKernelTiming(double *op_ms_total, unsigned int *op_launches)
{
    if (op_ms_total) *op_ms_total = g_op_ms;
    if (op_launches) *op_launches = g_op_count;
    return CUDAPCG_TRUE;
}

void hmgSolveHomogenization(void)
{
    if (!M) return;
    const double t_hmg = omp_get_wtime();
    appendf("Running homogenization...\n");
    if (!hmgPrepare()) return;
    hmgSolveResident(nullptr);
    hmgRelease();
    printf("Done.(%.2e s)\n", omp_get_wtime() - t_hmg);
}

void hmgPrintModelData(void)
{
    if (!M) return;
    static const char *an[3] = {"THERMAL_3D", "ELASTIC_3D", "FLUID_3D"};
    printf("*******************************************************\nMODEL DATA:\nAnalysis: %s\n", an[M->analysis]);
    printf("Mesh: [%u,%u,%u] elements, %zu DOFs\n", M->ex, M->ey, M->ez, M->ndof);
    if (M->analysis == HMG_FLUID)
        printf("  | porosity: %05.2f%%\n  | within pores: %u\n  | at interfaces: %u\n",
               100.0 * M->nporeelem / (double)M->nelem, M->nvel, M->nborder);
    printf("Solver tolerance: %.3e, max iterations: %u\n*******************************************************\n", M->tol,
           M->maxit);
}

void hmgPrintConstitutiveMtx(void)
{
    if (!M) return;
    const unsigned n = M->analysis == HMG_ELASTIC ? 6 : 3;
    printf("-------------------------------------------------------\nHomogenized Constitutive Matrix:\n");
    for (unsigned i = 0; i < n; i++) {
        printf(" ");
        for (unsigned j = 0; j < n; j++) printf(" %.8e", M->C[i * n + j]);
        printf("\n");
    }
    printf("-------------------------------------------------------\n");
    if (const double *a = hmgGetThermalExpansion()) {
        // symmetric 3x3 from Voigt xx, yy, zz, yz, xz, xy (femhmg_elastic_3D.c:1625-1630)
        printf("Effective thermal expansion:\n");
        printf("  %.8e  %.8e  %.8e\n", a[0], a[5], a[4]);
        printf("  %.8e  %.8e  %.8e\n", a[5], a[1], a[3]);
        printf("  %.8e  %.8e  %.8e\n", a[4], a[3], a[2]);
        printf("-------------------------------------------------------\n");
    }
}

var *hmgGetConstitutiveMtx(void) { return M ? M->C : nullptr; }
unsigned int hmgGetConstitutiveMtxDim(void) { return M ? M->C_dim : 0; }
var *hmgGetThermalExpansion(void) { return M && M->thermal_expansion && M->texp_ok ? M->texp : nullptr; }
unsigned int hmgGetThermalExpansionDim(void) { return hmgGetThermalExpansion() ? 6 : 0; }

logical hmgSaveConstitutiveMtx(const char *filename)
{
    if (!M || !filename) return HMG_FALSE;
    FILE *f = fopen(filename, "wb");
    if (!f) return HMG_FALSE;
    fwrite(M->C, sizeof(double), M->C_dim, f);
    if (hmgGetThermalExpansion()) fwrite(M->texp, sizeof(double), 6, f);
    fclose(f);
    return HMG_TRUE;
}

void hmgSaveFields(logical mustExport_flag, logical byElems_flag)
{
    (void)byElems_flag;
    if (M) M->save_fields = mustExport_flag > 0;
}

void hmgKeepTrackOfAnalysisReport(reportFlag_t flag) { if (M) M->report_on = flag > 0; }

reportFlag_t hmgPrintReport(const char *filename)
{
    if (!M || !M->report_on || !filename) return 0;
    FILE *f = fopen(filename, "w");
    if (!f) return 0;
    for (auto &s : M->report) fputs(s.c_str(), f);
    fclose(f);
    return 1;
}

cudapcgFlag_t hmgGetSolveStats(chfemSolveStats_t *stats, unsigned int *count)
{
    if (!M || !stats || !count) return CUDAPCG_FALSE;
    *count = (unsigned)M->stats.size();
    for (size_t i = 0; i < M->stats.size(); i++) stats[i] = M->stats[i];
    return CUDAPCG_TRUE;
}

logical hmgGetSizes(unsigned int *nelem, unsigned int *ndof, unsigned int *nvel, unsigned int *nborder)
{
    if (!M) return HMG_FALSE;
    if (nelem) *nelem = (unsigned)M->nelem;
    if (ndof) *ndof = (unsigned)M->ndof;
    if (nvel) *nvel = M->nvel;
    if (nborder) *nborder = M->nborder;
    return HMG_TRUE;
}

logical hmgGetLocalMatrices(double *Mtxs, double *CB)
{
    if (!M) return HMG_FALSE;
    if (Mtxs) memcpy(Mtxs, M->Mtxs.data(), M->Mtxs.size() * sizeof(double));
    if (CB) memcpy(CB, M->CB.data(), M->CB.size() * sizeof(double));
    return HMG_TRUE;
}

logical hmgGetMaps(uint16_t *keys, uint8_t *fluid_keys, uint32_t *dof_id, uint32_t *dof2node, uint8_t *border_keys)
{
    if (!M) return HMG_FALSE;
    if (keys && !M->keys.empty()) memcpy(keys, M->keys.data(), M->keys.size() * 2);
    if (fluid_keys && !M->fkeys.empty()) memcpy(fluid_keys, M->fkeys.data(), M->fkeys.size());
    if (dof_id && !M->dof_id.empty()) memcpy(dof_id, M->dof_id.data(), M->dof_id.size() * 4);
    if (dof2node && !M->dof2node.empty()) memcpy(dof2node, M->dof2node.data(), M->dof2node.size() * 4);
    if (border_keys && !M->bkeys.empty()) memcpy(border_keys, M->bkeys.data(), M->bkeys.size());
    return HMG_TRUE;
}

// RHS / C column on the device for the current model, through a temporary
// solver set up as hmgSolveHomogenization does.
logical hmgAssembleRHS(unsigned int dir, cudapcgVar_t *b)
{
    if (!M || !b) return HMG_FALSE;
    DeviceAux aux;
    if (!solver_setup(*M, aux)) return HMG_FALSE;
    bool ok = device_rhs(*M, aux, dir) &&
              hipMemcpy(b, rhs_buffer(), M->ndof * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
    free_aux(aux);
    cudapcgEnd();
    return ok ? HMG_TRUE : HMG_FALSE;
}

logical hmgUpdateC(unsigned int dir, const cudapcgVar_t *x, var *C)
{
    if (!M || !x || !C) return HMG_FALSE;
    DeviceAux aux;
    if (!solver_setup(*M, aux)) return HMG_FALSE;
    bool ok = hipMemcpy(solution_buffer(), x, M->ndof * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
              device_updateC(*M, aux, dir, solution_buffer());
    if (ok) memcpy(C, M->C, M->C_dim * sizeof(double));
    free_aux(aux);
    cudapcgEnd();
    return ok ? HMG_TRUE : HMG_FALSE;
}

// Allocation-free variant of hmgUpdateC for the model hmgPrepare made
// resident: the C column of load case dir from a device solution, with the
// reduction partials, sums and seam table in the static scratch array of
// noalloc.hip.  Same kernels and reduction order as hmgUpdateC.
logical hmgUpdateCNoAlloc(unsigned int dir, const double *x_dev, double *Ccol_dev)
{
    if (!M || !g_prepared) {
        fprintf(stderr, "chfem: hmgUpdateCNoAlloc needs a model made resident by hmgPrepare\n");
        return HMG_FALSE;
    }
    Model &m = *M;
    const unsigned nr = m.analysis == HMG_ELASTIC ? 6u : 3u;
    if (!x_dev || !Ccol_dev || dir >= nr) {
        fprintf(stderr, "chfem: hmgUpdateCNoAlloc: null vector or load case %u out of range\n", dir);
        return HMG_FALSE;
    }
    if (m.analysis != HMG_FLUID && m.nmat > (unsigned)NOALLOC_MAX_MAT) {
        fprintf(stderr, "chfem: hmgUpdateCNoAlloc: %u materials exceed the scratch limit of %d\n", m.nmat,
                NOALLOC_MAX_MAT);
        return HMG_FALSE;
    }
    const DeviceModel *dm = built_model();
    double *sc = noalloc_scratch();
    if (!dm || !sc) return HMG_FALSE;
    const int nb = reduce_blocks(m.analysis == HMG_FLUID ? std::max(m.nvel, 1u) : m.nelem);
    if ((size_t)nb * 6 > NOALLOC_PARTIALS) {
        fprintf(stderr, "chfem: hmgUpdateCNoAlloc: %d reduction blocks exceed the scratch limit\n", nb);
        return HMG_FALSE;
    }
    Reduce red;
    red.partials = sc;
    red.S = sc + NOALLOC_SUMS_OFF;
    red.fin = FIN_STORE;
    red.slot = 0;
    red.guard = 0;
    hipStream_t st = active_stream();
    if (m.analysis == HMG_FLUID) {
        launch_fluid_velocity_sum<double>(dm->op64, x_dev, &red, st);
        launch_noalloc_column(red.S, Ccol_dev, 3, m.elem_size * m.elem_size / (double)m.nelem, 0, st);
    } else {
        const LoadCase lc = load_case(m, dir);
        double seam[6 * NOALLOC_MAX_MAT];
        seam_table(m, lc, seam);
        launch_noalloc_seam(sc + NOALLOC_SEAM_OFF, seam, 6 * (int)m.nmat, st);
        launch_updateC(dm->op64, g_aux.CB, sc + NOALLOC_SEAM_OFF, lc, x_dev, &red, st);
        launch_noalloc_column(red.S, Ccol_dev, (int)nr, (double)m.nelem, 1, st);
    }
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        fprintf(stderr, "chfem: hmgUpdateCNoAlloc: %s\n", hipGetErrorString(e));
        return HMG_FALSE;
    }
    return HMG_TRUE;
}

// chfem/src/main.c:300-323
void initDefaultInput(chfemgpuInput_t *u)
{
    memset(u, 0, sizeof(*u));
    u->hmg_direction_flag = HOMOGENIZE_ALL;
    u->solver_flag = 0;
    u->preconditioner_flag = 1;
    u->parallel_flag = 0;
    u->stopcrit_flag = 0;
    u->poremap_flag = 1;
    u->xreduce_flag = 2;
}

// chfem/src/main.c:72-152
int runAnalysis(chfemgpuInput_t *u)
{
    if (u->writeReport_flag) hmgKeepTrackOfAnalysisReport(1);
    hmgSaveFields(u->exportFields_flag, u->fieldsByElem_flag);
    hmgSetParallelStrategyFlag(u->parallel_flag);
    if (!hmgSetPoreMappingStrategy(u->poremap_flag)) {
        printf("ERROR: Failed to set \"poremap_flag\"=%d.\nProcess aborted.\n", u->poremap_flag);
        hmgEnd();
        return -1;
    }
    hmgSetSolverFlag(u->solver_flag);
    hmgSetPreConditionerFlag(u->preconditioner_flag);
    hmgSetXReduceFlag(u->xreduce_flag);
    hmgSetHomogenizationFlag(u->hmg_direction_flag);
    hmgPrintModelData();
    hmgSetStoppingCriteria(u->stopcrit_flag);
    if (u->importX_flag) hmgImportX(u->x0_file);
    else hmgFindInitialGuesses(u->num_of_recursions_initguess);
    hmgExportX(u->exportX_flag);
    hmgSolveHomogenization();
    hmgPrintConstitutiveMtx();
    hmgPrintReport(u->report_file);
    if (u->save2binary_flag) hmgSaveConstitutiveMtx(u->binary_file);
    const double *C = hmgGetConstitutiveMtx();
    const unsigned cd = hmgGetConstitutiveMtxDim();
    for (unsigned i = 0; i < cd; i++) u->eff_coeff[i] = C[i];
    u->num_of_thermal_expansion_coeffs = 0;
    if (const double *a = hmgGetThermalExpansion()) {
        u->num_of_thermal_expansion_coeffs = hmgGetThermalExpansionDim();
        for (unsigned i = 0; i < u->num_of_thermal_expansion_coeffs; i++) u->eff_coeff[cd + i] = a[i];
    }
    hmgEnd();
    printf("#######################################################\n");
    return 0;
}

// chfem/src/chfem_wrapper.c:7-89, as a plain C ABI
int chfem_run(const uint8_t *img, size_t img_size, const char *nf_file, int analysis, int direction, int solver,
              int precondition, int output_fields, double *eff_out, unsigned *n_out, unsigned *alpha_out)
{
    (void)img_size;
    chfemgpuInput_t u;
    initDefaultInput(&u);
    u.hmg_direction_flag = (unsigned)direction;
    u.solver_flag = (unsigned)solver;
    u.preconditioner_flag = (unsigned)precondition;
    u.exportFields_flag = (unsigned char)output_fields;
    if (!hmgInit((char *)nf_file, nullptr, nullptr, (uint8_t *)img)) {
        printf("Failed to properly read input files.\nProcess aborted.\n");
        return -1;
    }
    if (runAnalysis(&u) != 0) return -1;
    const unsigned n = (analysis == 0 || analysis == 2) ? 3 : 6;
    for (unsigned i = 0; i < n * n + u.num_of_thermal_expansion_coeffs; i++) eff_out[i] = u.eff_coeff[i];
    if (n_out) *n_out = n;
    if (alpha_out) *alpha_out = u.num_of_thermal_expansion_coeffs;
    return 0;
}

}  // extern "C"

// ===========================================================================
// Batched solver: every load case of many samples (one analysis type, any
// sizes) advanced together by the batch kernels (batch.hip).  Mixed
// precision: FP32 CG with FP64 scalars and FP64 residual replacement per
// system (the scheme of the single-system mixed solver).
// ===========================================================================
namespace {

struct BatchSample {
    Model *m = nullptr;
    uint16_t *keys = nullptr;
    void *mat32 = nullptr, *mat64 = nullptr;
    float *kinv32 = nullptr;
    double *kinv64 = nullptr;
    double *K64 = nullptr, *CB64 = nullptr, *seam = nullptr;
    std::vector<chfemSolveStats_t> stats;
    double C[36];
};

struct BatchRun {
    std::vector<BatchSample> samples;
    int analysis = -1;
    double seconds = 0;
    // learned initial guess for every system (nullptr: zero guess)
    void *net = nullptr;
};

template <typename T> void bfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <typename T> bool bupload(T *&dst, const T *src, size_t n)
{
    if (hipMalloc(&dst, n * sizeof(T)) != hipSuccess) return false;
    return hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

void free_sample(BatchSample &b)
{
    bfree(b.keys); bfree(b.mat32); bfree(b.mat64); bfree(b.kinv32); bfree(b.kinv64);
    bfree(b.K64); bfree(b.CB64); bfree(b.seam);
    delete b.m;
    b.m = nullptr;
}

OpDesc sample_op(const BatchSample &b)
{
    const Model &m = *b.m;
    OpDesc o;
    memset(&o, 0, sizeof(o));
    o.physics = m.analysis == HMG_ELASTIC ? PHY_ELASTIC : PHY_THERMAL;
    o.ny = (int)m.ey; o.nx = (int)m.ex; o.nz = (int)m.ez;
    o.N = (uint32_t)m.nelem;
    o.ndof = m.ndof;
    o.keys = b.keys;
    o.K = b.K64;
    o.nmat = (int)m.nmat;
    return o;
}

bool upload_sample(BatchSample &b)
{
    const Model &m = *b.m;
    const int nd = m.analysis == HMG_ELASTIC ? 3 : 1;
    const std::vector<double> K = dense_K(m);
    if (!bupload(b.keys, m.keys.data(), m.keys.size()) || !bupload(b.K64, K.data(), K.size()) ||
        !bupload(b.CB64, m.CB.data(), m.CB.size()))
        return false;
    if (hipMalloc(&b.seam, 6 * m.nmat * sizeof(double)) != hipSuccess) return false;
    std::vector<double> t64;
    std::vector<float> t32;
    voxel::inverse_diagonal_table(K.data(), m.nmat, nd, t64, t32);
    if (!bupload(b.kinv64, t64.data(), t64.size()) || !bupload(b.kinv32, t32.data(), t32.size())) return false;
    if (m.analysis == HMG_ELASTIC) {
        double lm[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        float lmf[8];
        for (unsigned t = 0; t < m.nmat; t++)
            if (!voxel::isotropic_params(K.data() + 576 * t, lm[2 * t], lm[2 * t + 1])) return false;
        for (int i = 0; i < 8; i++) lmf[i] = (float)lm[i];
        double *d64 = nullptr;
        float *d32 = nullptr;
        if (!bupload(d64, lm, 8) || !bupload(d32, lmf, 8)) return false;
        b.mat64 = d64;
        b.mat32 = d32;
    } else {
        // conduction: {k, 0} per material for the Walsh-basis operator
        double kk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        float kf[8];
        for (unsigned t = 0; t < m.nmat; t++)
            if (!voxel::conduction_param(K.data() + 64 * t, kk[2 * t])) return false;
        for (int i = 0; i < 8; i++) kf[i] = (float)kk[i];
        double *d64 = nullptr;
        float *d32 = nullptr;
        if (!bupload(d64, kk, 8) || !bupload(d32, kf, 8)) return false;
        b.mat64 = d64;
        b.mat32 = d32;
    }
    return true;
}

std::vector<double> seam_values(const Model &m, const LoadCase &lc)
{
    const int nd = m.analysis == HMG_ELASTIC ? 3 : 1, ke = 8 * nd, nr = nd == 3 ? 6 : 3;
    std::vector<double> seam(6 * m.nmat, 0.0);
    for (unsigned t = 0; t < m.nmat; t++)
        for (int row = 0; row < nr; row++) {
            double sum = 0;
            for (int k = 0; k < 4; k++) sum += m.CB[(size_t)t * nr * ke + row * ke + nd * lc.S[k] + lc.comp];
            seam[t * 6 + row] = sum * lc.L;
        }
    return seam;
}

}  // namespace

extern "C" {

void *chfemBatchCreate(void) { return new BatchRun(); }

void chfemBatchDestroy(void *h)
{
    BatchRun *b = (BatchRun *)h;
    if (!b) return;
    for (BatchSample &s : b->samples) free_sample(s);
    delete b;
}

int chfemBatchAddSample(void *h, const char *nf_file, const uint8_t *img)
{
    BatchRun *b = (BatchRun *)h;
    if (!b || !nf_file || !img) return 0;
    Model *m = new Model();
    memset(m->props, 0, sizeof(m->props));
    memset(m->alpha, 0, sizeof(m->alpha));
    memset(m->props_keys, 0, sizeof(m->props_keys));
    memset(m->C, 0, sizeof(m->C));
    m->nf = nf_file;
    bool ok = read_data(*m, nf_file) && init_sizes(*m);
    if (ok && m->analysis == HMG_FLUID) {
        printf("ERROR: the batched solver takes thermal and elastic models.\n");
        ok = false;
    }
    if (ok && m->nmat > 4) {
        printf("ERROR: the batched solver takes at most 4 materials per sample.\n");
        ok = false;
    }
    if (ok && b->analysis >= 0 && m->analysis != b->analysis) {
        printf("ERROR: all samples of a batch share one analysis type.\n");
        ok = false;
    }
    ok = ok && read_image(*m, img);
    if (!ok) {
        delete m;
        return 0;
    }
    build_keys(*m);
    assemble_local(*m);
    b->analysis = m->analysis;
    BatchSample s;
    s.m = m;
    if (!upload_sample(s)) {
        free_sample(s);
        return 0;
    }
    b->samples.push_back(s);
    return 1;
}

}  // extern "C"

// Learned initial guesses of a batch (see chfemBatchSetInitialGuessNetwork):
// x64 of every system of a sample = ramp + the network's prediction for its
// load case, then one FP64 replacement step (r64 -= A x64, r.M^-1 r per
// system).  A system whose guess leaves a larger residual than b goes back
// to x = 0, r = b.  Leaves R[SB_DELTA] = r.M^-1 r of the guess and x0used
// set where a guess was kept.
static bool batch_guess(BatchRun *B, const batch::Batch &bt, const std::vector<batch::Sys> &sys,
                        const std::vector<std::pair<int, int>> &owner, int nd, double *x64, double *r64, double *q64,
                        float *r32, std::vector<double> &Sh, std::vector<double> &x0res, std::vector<int> &x0used,
                        hipStream_t st)
{
    using namespace batch;
    const int nsys = (int)sys.size();
    auto R = [&](int s, int slot) -> double & { return Sh[(size_t)s * SB_STRIDE + slot]; };
    std::vector<int> has(nsys, 0);
    for (size_t si = 0; si < B->samples.size(); si++) {
        const Model &m = *B->samples[si].m;
        int gC = 0;
        const float *gy = guess_predict_device(B->net, m.elem_map.data(), m.props, m.nmat, m.ex, m.ey, m.ez, &gC);
        if (!gy || hipStreamSynchronize(guess_stream(B->net)) != hipSuccess) continue;
        double *xs[6] = {};
        for (int s = 0; s < nsys; s++) {
            if (owner[s].first != (int)si || fabs(R(s, SB_DELTA0)) < 1e-12) continue;
            xs[owner[s].second] = x64 + sys[s].off;
            has[s] = 1;
        }
        nn::guess_to_x0_all(gy, gC, nd, nn::Grid3{(int)m.ey, (int)m.ex, (int)m.ez}, 1.0f, xs, st);
        // the next prediction reuses the network's output buffer
        if (hipStreamSynchronize(st) != hipSuccess) return false;
    }
    for (int s = 0; s < nsys; s++) R(s, SB_REPL) = has[s];
    if (hipMemcpyAsync((double *)bt.Sb, Sh.data(), Sh.size() * 8, hipMemcpyHostToDevice, st) != hipSuccess) return false;
    apply<double>(bt, x64, q64, OP_REPL, st);
    resid(bt, r64, q64, r32, st);
    finish(bt, FIN_B_TRUE, st);
    if (hipMemcpyAsync(Sh.data(), bt.Sb, Sh.size() * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return false;
    for (int s = 0; s < nsys; s++) {
        if (!has[s]) continue;
        R(s, SB_REPL) = 0.0;
        const double dl = R(s, SB_DELTA), d0 = R(s, SB_DELTA0);
        x0res[s] = sqrt(fabs(dl) / fabs(d0));
        if (fabs(dl) > fabs(d0)) {
            const size_t len = (size_t)nd * sys[s].N;
            launch_add<double>(r64 + sys[s].off, q64 + sys[s].off, len, st);
            launch_fill<double>(x64 + sys[s].off, 0.0, len, st);
            R(s, SB_DELTA) = d0;
        } else {
            x0used[s] = 1;
        }
    }
    return hipMemcpyAsync((double *)bt.Sb, Sh.data(), Sh.size() * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
           hipStreamSynchronize(st) == hipSuccess;
}

extern "C" {

int chfemBatchSetInitialGuessNetwork(void *h, void *net)
{
    BatchRun *B = (BatchRun *)h;
    if (!B) return 0;
    if (net && guess_physics(net) != (B->analysis == HMG_ELASTIC ? 1 : 0) && B->analysis >= 0) {
        fprintf(stderr, "chfem: the guess network's physics differs from the batch's analysis\n");
        return 0;
    }
    B->net = net;
    return 1;
}

// The operator's tile table in dispatch order.  The kernels give block b
// table entry start(b % 8) + b / 8 (the hardware sends block b to XCD
// b % 8), so XCD x runs the contiguous range [start(x), start(x) + n_x),
// n_x = T / 8 (+1 for x < T % 8).  That range is made of one contiguous
// eighth of every system's tiles (the remainders of a system's split go to
// the XCDs in rotation, continuing across systems, which gives exactly those
// n_x): tiles that share halo nodes meet in one XCD's L2, and every XCD has
// work for as long as any system iterates.  tiles comes in system-contiguous
// order and keeps that position in Tile::slot.
static std::vector<batch::Tile> xcd_table(const std::vector<batch::Sys> &sys, const std::vector<batch::Tile> &tiles)
{
    std::vector<batch::Tile> out;
    out.reserve(tiles.size());
    std::vector<int> first(sys.size() * 8), count(sys.size() * 8);
    int rot = 0;
    for (size_t s = 0; s < sys.size(); s++) {
        const int n = sys[s].tb1 - sys[s].tb0, base = n / 8, extra = n % 8;
        int at = sys[s].tb0;
        for (int x = 0; x < 8; x++) {
            const int c = base + ((x - rot + 8) % 8 < extra ? 1 : 0);
            first[s * 8 + x] = at;
            count[s * 8 + x] = c;
            at += c;
        }
        rot = (rot + extra) % 8;
    }
    for (int x = 0; x < 8; x++)
        for (size_t s = 0; s < sys.size(); s++)
            for (int i = 0; i < count[s * 8 + x]; i++) out.push_back(tiles[first[s * 8 + x] + i]);
    return out;
}

int chfemBatchSolve(void *h, double tol, unsigned int maxit)
{
    using namespace batch;
    BatchRun *B = (BatchRun *)h;
    if (!B || B->samples.empty()) return 0;
    const int nd = B->analysis == HMG_ELASTIC ? 3 : 1, nloads = nd == 3 ? 6 : 3;
    const double drop = 1e-2;
    // systems and work tables
    std::vector<Sys> sys;
    std::vector<Tile> tiles;
    std::vector<Chunk> chunks;
    std::vector<std::pair<int, int>> owner;  // (sample, load) of each system
    size_t off = 0;
    for (size_t si = 0; si < B->samples.size(); si++) {
        const Model &m = *B->samples[si].m;
        for (int d = 0; d < nloads; d++) {
            Sys S;
            memset(&S, 0, sizeof(S));
            S.ny = (int)m.ey; S.nx = (int)m.ex; S.nz = (int)m.ez;
            S.LZ = TILE_LZ;
            S.N = (uint32_t)m.nelem;
            S.off = off;
            S.keys = B->samples[si].keys;
            S.mat32 = B->samples[si].mat32; S.mat64 = B->samples[si].mat64;
            S.kinv32 = B->samples[si].kinv32; S.kinv64 = B->samples[si].kinv64;
            const int sid = (int)sys.size();
            S.tb0 = (int)tiles.size();
            {
                const int gx = (S.ny + TILE_ROWS - 1) / TILE_ROWS, gy = (S.nx + TILE_C - 2) / (TILE_C - 1);
                const int gz = (S.nz + S.LZ - 1) / S.LZ;
                for (int z = 0; z < gz; z++)
                    for (int y = 0; y < gy; y++)
                        for (int x = 0; x < gx; x++) tiles.push_back(Tile{sid, x, y, z, (int)tiles.size()});
            }
            S.tb1 = (int)tiles.size();
            S.vb0 = (int)chunks.size();
            for (uint32_t i0 = 0; i0 < S.N; i0 += 256) chunks.push_back(Chunk{sid, i0, std::min(S.N, i0 + 256u)});
            S.vb1 = (int)chunks.size();
            sys.push_back(S);
            owner.push_back({(int)si, d});
            off += (size_t)nd * m.nelem;
        }
    }
    const size_t total = off;
    const int nsys = (int)sys.size();
    tiles = xcd_table(sys, tiles);
    hipStream_t st = nullptr;
    if (hipStreamCreate(&st) != hipSuccess) return 0;
    Sys *dsys = nullptr;
    Tile *dtiles = nullptr;
    Chunk *dchunks = nullptr;
    double *Sb = nullptr, *part = nullptr, *x64 = nullptr, *r64 = nullptr, *e64 = nullptr, *q64 = nullptr;
    double *rpart = nullptr, *rS = nullptr;
    float *x32 = nullptr, *r32 = nullptr, *d32 = nullptr, *q32 = nullptr, *ring = nullptr;
    bool ok = bupload(dsys, sys.data(), sys.size()) && bupload(dtiles, tiles.data(), tiles.size()) &&
              bupload(dchunks, chunks.data(), chunks.size()) &&
              hipMalloc(&Sb, sizeof(double) * nsys * SB_STRIDE) == hipSuccess &&
              hipMalloc(&part, sizeof(double) * std::max(4 * tiles.size(), chunks.size())) == hipSuccess &&
              hipMalloc(&x64, total * 8) == hipSuccess && hipMalloc(&r64, total * 8) == hipSuccess &&
              hipMalloc(&e64, total * 8) == hipSuccess && hipMalloc(&q64, total * 8) == hipSuccess &&
              hipMalloc(&x32, total * 4) == hipSuccess && hipMalloc(&r32, total * 4) == hipSuccess &&
              hipMalloc(&q32, total * 4) == hipSuccess &&
              hipMalloc(&rpart, sizeof(double) * MAX_RED_GRID * 8) == hipSuccess &&
              hipMalloc(&rS, sizeof(double) * S_COUNT) == hipSuccess;
    // Fused loop on z (r32 holds z): directions in a ring of w slots (16, or
    // fewer if they do not fit; w divides the 16 steps between host
    // synchronisations).  Without room for two slots: the separate
    // direction, operator and update kernels on r.
    int w = 0;
    // CHFEM_BATCH_RING caps the width (16, 8, 4 or 2: a width divides the 16
    // steps between host synchronisations, where the window is flushed)
    int wmax = SB_RING_MAX;
    if (const char *re = getenv("CHFEM_BATCH_RING"))
        while (wmax > 2 && wmax > atoi(re)) wmax /= 2;
    for (int cand = wmax; ok && cand >= 2 && !w; cand /= 2)
        if (hipMalloc(&ring, total * 4 * (size_t)cand) == hipSuccess) w = cand;
        else (void)hipGetLastError();
    const bool zform = w >= 2;
    if (zform) ok = ok && hipMemsetAsync(ring, 0, total * 4 * (size_t)w, st) == hipSuccess;
    else ok = ok && hipMalloc(&d32, total * 4) == hipSuccess;
    // single-kernel steps (the default; CHFEM_BATCH_FORM=1 keeps the
    // operator with the direction update and a separate z update): a second
    // z and q buffer, which z and q alternate with by each system's
    // iteration parity
    float *z2 = nullptr, *q2 = nullptr;
    const char *bf = getenv("CHFEM_BATCH_FORM");
    bool cg1 = zform && !(bf && atoi(bf) == 1);
    if (ok && cg1) {
        if (hipMalloc(&z2, total * 4) != hipSuccess || hipMalloc(&q2, total * 4) != hipSuccess) {
            (void)hipGetLastError();
            bfree(z2);
            bfree(q2);
            cg1 = false;
        }
    }
    Batch bt;
    bt.nsys = nsys; bt.nd = nd; bt.sys = dsys; bt.tiles = dtiles; bt.ntiles = (int)tiles.size();
    bt.chunks = dchunks; bt.nchunks = (int)chunks.size(); bt.Sb = Sb; bt.part = part;
    std::vector<double> Sh((size_t)nsys * SB_STRIDE, 0.0);
    auto down = [&]() {
        return hipMemcpyAsync(Sh.data(), Sb, Sh.size() * 8, hipMemcpyDeviceToHost, st) == hipSuccess &&
               hipStreamSynchronize(st) == hipSuccess;
    };
    auto up = [&]() {
        return hipMemcpyAsync(Sb, Sh.data(), Sh.size() * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
               hipStreamSynchronize(st) == hipSuccess;
    };
    auto R = [&](int s, int slot) -> double & { return Sh[(size_t)s * SB_STRIDE + slot]; };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ok = ok && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    std::vector<double> d0v(nsys, 0.0), delta_rel(nsys, 0.0), x0res(nsys, 0.0);
    std::vector<int> x0used(nsys, 0);
    std::vector<unsigned> repl(nsys, 0);
    std::vector<int> conv(nsys, 0);
    if (ok) {
        hipEventRecord(e0, st);
        // right-hand sides (FP64) and x = 0
        for (int s = 0; s < nsys; s++) {
            BatchSample &bs = B->samples[owner[s].first];
            const LoadCase lc = load_case(*bs.m, (unsigned)owner[s].second);
            launch_rhs(sample_op(bs), bs.K64, lc, owner[s].second, r64 + sys[s].off, st);
        }
        launch_fill<double>(x64, 0.0, total, st);
        delta0(bt, r64, st);
        for (int s = 0; s < nsys; s++) R(s, SB_ACTIVE) = 0.0;
        ok = up();
        finish(bt, FIN_B_DELTA0, st);
        ok = ok && down();
        // learned initial guess: x64 = ramp + prediction for every system of
        // a sample the network can predict; the FP64 residual b - A x0 then
        // replaces b, unless it is larger than b's (the zero guess)
        if (ok && B->net) ok = batch_guess(B, bt, sys, owner, nd, x64, r64, q64, r32, Sh, x0res, x0used, st);
        // fresh FP32 start of every system with a non-zero right-hand side
        for (int s = 0; s < nsys; s++) {
            d0v[s] = R(s, SB_DELTA0);
            delta_rel[s] = x0used[s] ? R(s, SB_DELTA) : d0v[s];
            const bool live = fabs(d0v[s]) >= 1e-12;
            conv[s] = live ? 0 : 1;
            R(s, SB_ACTIVE) = live;
            R(s, SB_START) = live;
            R(s, SB_FRESH) = 1.0;
            R(s, SB_DONE) = 0.0;
            R(s, SB_REPL) = 0.0;
            R(s, SB_ITER) = 0.0;
            R(s, SB_BASE) = 0.0;
            R(s, SB_THRESH) = std::max(tol * tol * fabs(d0v[s]), drop * fabs(delta_rel[s]));
            if (live && x0used[s] && fabs(delta_rel[s]) <= tol * tol * fabs(d0v[s])) {
                // the guess already meets the tolerance
                conv[s] = 1;
                R(s, SB_ACTIVE) = 0.0;
                R(s, SB_START) = 0.0;
            }
        }
        ok = ok && up();
        if (zform) start_z(bt, r64, r32, x32, st, cg1 ? z2 : nullptr);
        else start(bt, r64, r32, d32, x32, st);
        finish(bt, FIN_B_START, st);
        ok = ok && down();
        for (int s = 0; s < nsys; s++) R(s, SB_START) = 0.0;
        ok = ok && up();
    }
    const unsigned nb = 16;
    while (ok) {
        int live = 0;
        for (int s = 0; s < nsys; s++) live += R(s, SB_ACTIVE) != 0.0;
        if (!live) break;
        for (unsigned k = 0; k < nb; k++) {
            if (cg1) {
                apply_cg1(bt, r32, z2, q32, q2, ring, total, w, st);
                if ((k + 1) % (unsigned)w == 0) ring_flush(bt, x32, ring, total, w, st);
                continue;
            }
            if (zform) {
                apply_dir(bt, r32, ring, total, w, q32, st);
                finish_alpha(bt, w, st);
                zupdate(bt, r32, q32, st);
                finish(bt, FIN_B_DELTA, st);
                if ((k + 1) % (unsigned)w == 0) ring_flush(bt, x32, ring, total, w, st);
                continue;
            }
            dir(bt, d32, r32, st);
            apply<float>(bt, d32, q32, OP_ITER, st);
            finish(bt, FIN_B_ALPHA, st);
            update(bt, x32, r32, d32, q32, st);
            finish(bt, FIN_B_DELTA, st);
        }
        if (!(ok = down())) break;
        int nrepl = 0;
        for (int s = 0; s < nsys; s++) {
            R(s, SB_REPL) = 0.0;
            if (R(s, SB_ACTIVE) == 0.0) continue;
            if (R(s, SB_DONE) != 0.0 || R(s, SB_ITER) >= maxit) {
                R(s, SB_REPL) = 1.0;
                nrepl++;
            }
        }
        if (!nrepl) continue;
        if (!(ok = up())) break;
        // residual replacement of the selected systems
        flush(bt, x64, e64, x32, st);
        apply<double>(bt, e64, q64, OP_REPL, st);
        if (zform) resid_z(bt, r64, q64, r32, st, cg1 ? z2 : nullptr);
        else resid(bt, r64, q64, r32, st);
        finish(bt, FIN_B_TRUE, st);
        if (!(ok = down())) break;
        for (int s = 0; s < nsys; s++) {
            if (R(s, SB_REPL) == 0.0) continue;
            R(s, SB_REPL) = 0.0;
            repl[s]++;
            const double dl = R(s, SB_DELTA);
            if (fabs(dl) <= tol * tol * fabs(d0v[s])) {
                conv[s] = 1;
                R(s, SB_ACTIVE) = 0.0;
            } else if (R(s, SB_ITER) >= maxit || !(fabs(dl) < fabs(delta_rel[s]))) {
                R(s, SB_ACTIVE) = 0.0;  // iteration cap, or no progress left in FP32
            } else {
                delta_rel[s] = dl;
                R(s, SB_DONE) = 0.0;
                R(s, SB_BASE) = R(s, SB_ITER);
                R(s, SB_THRESH) = std::max(tol * tol * fabs(d0v[s]), drop * fabs(dl));
            }
        }
        ok = up();
    }
    float ms = 0;
    if (ok) {
        hipEventRecord(e1, st);
        ok = hipEventSynchronize(e1) == hipSuccess;
        hipEventElapsedTime(&ms, e0, e1);
    }
    B->seconds = ms * 1e-3;
    // effective tensors (FP64, on the device)
    for (BatchSample &bs : B->samples) {
        memset(bs.C, 0, sizeof(bs.C));
        bs.stats.assign(nloads, chfemSolveStats_t());
    }
    std::vector<double> Shr(S_COUNT);
    for (int s = 0; ok && s < nsys; s++) {
        BatchSample &bs = B->samples[owner[s].first];
        const Model &m = *bs.m;
        const int d = owner[s].second;
        const LoadCase lc = load_case(m, (unsigned)d);
        const std::vector<double> seam = seam_values(m, lc);
        ok = hipMemcpy(bs.seam, seam.data(), seam.size() * 8, hipMemcpyHostToDevice) == hipSuccess;
        Reduce red = {rpart, rS, FIN_STORE, S_TMP0, 0};
        launch_updateC(sample_op(bs), bs.CB64, bs.seam, lc, x64 + sys[s].off, &red, st);
        ok = ok && hipMemcpyAsync(Shr.data(), rS, S_COUNT * 8, hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
        const int nr = nd == 3 ? 6 : 3;
        for (int row = 0; row < nr; row++) bs.C[row * nr + d] = Shr[S_TMP0 + row] / (double)m.nelem;
        chfemSolveStats_t &stt = bs.stats[d];
        stt.iterations = (unsigned)R(s, SB_ITER);
        stt.sweeps = repl[s];
        stt.solve_seconds = B->seconds;
        stt.residual = sqrt(fabs(R(s, SB_DELTA)) / std::max(fabs(d0v[s]), 1e-300));
        stt.x0_residual = x0res[s];
        stt.x0_used = (unsigned)x0used[s];
    }
    int all = 1;
    for (int s = 0; s < nsys; s++) all &= conv[s];
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    bfree(dsys); bfree(dtiles); bfree(dchunks); bfree(Sb); bfree(part); bfree(x64); bfree(r64); bfree(e64);
    bfree(q64); bfree(x32); bfree(r32); bfree(d32); bfree(q32); bfree(ring); bfree(rpart); bfree(rS);
    bfree(z2); bfree(q2);
    (void)hipStreamDestroy(st);
    if (!ok) return 0;
    return all ? 1 : 2;
}

int chfemBatchGetC(void *h, unsigned int sample, double *C)
{
    BatchRun *b = (BatchRun *)h;
    if (!b || sample >= b->samples.size() || !C) return 0;
    const unsigned n = b->analysis == HMG_ELASTIC ? 6 : 3;
    memcpy(C, b->samples[sample].C, n * n * sizeof(double));
    return 1;
}

int chfemBatchGetStats(void *h, unsigned int sample, chfemSolveStats_t *stats, unsigned int *count)
{
    BatchRun *b = (BatchRun *)h;
    if (!b || sample >= b->samples.size()) return 0;
    const std::vector<chfemSolveStats_t> &v = b->samples[sample].stats;
    if (count) *count = (unsigned)v.size();
    if (stats)
        for (size_t i = 0; i < v.size(); i++) stats[i] = v[i];
    return 1;
}

double chfemBatchSeconds(void *h) { return h ? ((BatchRun *)h)->seconds : -1.0; }

}  // extern "C"
