// C-ABI of the eqdyna_amd time step: takes the reference's globalvar arrays,
// builds the device layout (see eqd_device.h), runs steps, returns state.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "eqd_internal.h"
#include "eqd_precompute.h"

using namespace eqd;

namespace {

struct FieldSpec {
    bool is_int;
    int64_t count;
};

}  // namespace

static bool field_spec(const eqd_scalars &s, int f, FieldSpec &o)
{
    const int64_t nn = s.nn, ne = s.ne, neq = s.neq, nf = s.nf;
    const int64_t non = s.non > 0 ? s.non : 1;
    switch (f) {
    case EQD_MESHCOOR: o = {false, 3 * nn}; return true;
    case EQD_EQNUMINDEXARR: o = {true, s.size_eqidx}; return true;
    case EQD_EQNUMSTARTINDEXLOC: o = {true, nn}; return true;
    case EQD_NUMOFDOFPERNODEARR: o = {true, nn}; return true;
    case EQD_NODEELEMIDRELATION: o = {true, 8 * ne}; return true;
    case EQD_ELEMTYPEARR: o = {true, ne}; return true;
    case EQD_MAT: o = {false, 5 * ne}; return true;
    case EQD_ELESHP: o = {false, 24 * ne}; return true;
    case EQD_ELEDET: o = {false, ne}; return true;
    case EQD_ELEMASS: o = {false, 24 * ne}; return true;
    case EQD_SS: o = {false, 6 * ne}; return true;
    case EQD_PHI: o = {false, 32 * ne}; return true;
    case EQD_STRESSCOMPINDEXARR: o = {true, ne}; return true;
    case EQD_ELEPOREP: o = {false, ne}; return true;
    case EQD_NODALMASSARR: o = {false, neq}; return true;
    case EQD_FNMS: o = {false, nn}; return true;
    case EQD_NSMP: o = {true, 2 * nf}; return true;
    case EQD_UN: case EQD_US: case EQD_UD: o = {false, 3 * nf}; return true;
    case EQD_ARN: o = {false, nf}; return true;
    case EQD_ANONFS: o = {true, 3 * (int64_t)s.non}; return true;
    case EQD_IDHIST: o = {true, 3 * 6 * (int64_t)s.noff}; return true;
    case EQD_STRESSARR: o = {false, s.size_stress}; return true;
    case EQD_PSTRAIN: o = {false, ne}; return true;
    case EQD_V1: o = {false, neq}; return true;
    case EQD_VELARR: case EQD_DISPARR: o = {false, 3 * nn}; return true;
    case EQD_NODALFORCEARR: o = {false, neq}; return true;
    case EQD_FRIC: o = {false, 100 * nf}; return true;
    case EQD_FNFT: o = {false, nf}; return true;
    case EQD_ONFAULTHIST: o = {false, 12 * (int64_t)s.nstep * non}; return true;
    case EQD_OFFFAULTHIST: o = {false, (6 * (int64_t)s.noff + 1) * s.nstep}; return true;
    default: return false;
    }
}

template <typename T>
static int dalloc(eqd_handle *h, T **p, int64_t count)
{
    *p = nullptr;
    if (count <= 0) return 0;
    void *q = nullptr;
    const size_t b = sizeof(T) * (size_t)count;
    if (hipMalloc(&q, b) != hipSuccess) return fail(h, EQD_EOOM, "hipMalloc of %zu bytes failed", b);
    h->allocs.push_back(q);
    h->bytes += (int64_t)b;
    *p = (T *)q;
    return 0;
}

template <typename T>
static int dupload(eqd_handle *h, T **p, const std::vector<T> &v)
{
    int rc = dalloc(h, p, (int64_t)v.size());
    if (rc) return rc;
    if (!v.empty()) HIPCHK(h, hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return 0;
}

// Component-major [w][n] values to the kernels' AoSoA blocks (blk_at<w>),
// zero padded to whole blocks of 64 items.
template <typename T, typename S>
static std::vector<T> to_blocks(const std::vector<S> &cm, int64_t n, int w)
{
    std::vector<T> b((size_t)((n + 63) / 64 * 64 * w), T(0));
    for (int64_t t = 0; t < n; t++)
        for (int k = 0; k < w; k++) b[(size_t)(((t >> 6) * w + k) * 64 + (t & 63))] = (T)cm[(size_t)k * n + t];
    return b;
}
static inline int64_t blocked_count(int64_t n, int w) { return (n + 63) / 64 * 64 * w; }
static inline int64_t blocked_index(int64_t t, int k, int w) { return ((t >> 6) * w + k) * 64 + (t & 63); }

// bfloat16 bits of a finite float (round to nearest even) and back, as the
// kernels store the FP32 bulk mode's displacement residual
static inline uint16_t bf16_bits(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
static inline float bf16_value(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// comdampv, src/comdampv.f90:1-120 (region 14 compares y with xmax0 as the
// reference does).  Returns d_x, d_y, d_z.
static void pml_node_damping(const eqd_scalars &s, double x2, double y, double z, double d[3])
{
    const double xmax0 = s.PMLb[0], xmin0 = s.PMLb[1], ymax0 = s.PMLb[2], ymin0 = s.PMLb[3], zmin0 = s.PMLb[4];
    double damp[3];
    damp[2] = z <= zmin0 ? std::fabs(z - zmin0) : 0.0;
    if (x2 >= xmax0 && y >= ymax0) { damp[0] = std::fabs(x2 - xmax0); damp[1] = std::fabs(y - ymax0); }
    else if (x2 >= xmax0 && y <= ymin0) { damp[0] = std::fabs(x2 - xmax0); damp[1] = std::fabs(y - ymin0); }
    else if (x2 <= xmin0 && y <= ymin0) { damp[0] = std::fabs(x2 - xmin0); damp[1] = std::fabs(y - ymin0); }
    else if (x2 <= xmin0 && y >= xmax0) { damp[0] = std::fabs(x2 - xmin0); damp[1] = std::fabs(y - ymax0); }
    else if (x2 >= xmax0 && y > ymin0 && y < ymax0) { damp[0] = std::fabs(x2 - xmax0); damp[1] = 0.0; }
    else if (y <= ymin0 && x2 > xmin0 && x2 < xmax0) { damp[0] = 0.0; damp[1] = std::fabs(y - ymin0); }
    else if (x2 <= xmin0 && y > ymin0 && y < ymax0) { damp[0] = std::fabs(x2 - xmin0); damp[1] = 0.0; }
    else if (y >= ymax0 && x2 > xmin0 && x2 < xmax0) { damp[0] = 0.0; damp[1] = std::fabs(y - ymax0); }
    else { damp[0] = 0.0; damp[1] = 0.0; }
    for (int i = 0; i < 3; i++) {
        const double delta = (double)s.nPML * s.PMLb[5 + i];
        d[i] = 3.0 * s.vmaxPML / 2.0 / delta * std::log(1.0 / s.R) * std::pow(damp[i] / delta, 2.0);
    }
}

// element-centre damping of calcPMLElemKU, src/assembleGlobalKU.f90:130-213
static void pml_elem_damping(const eqd_scalars &s, const double xc[3], double d[3])
{
    const double xmax2 = s.PMLb[0], xmin2 = s.PMLb[1], ymax2 = s.PMLb[2], ymin2 = s.PMLb[3], zmin2 = s.PMLb[4];
    double damps[3] = {0.0, 0.0, 0.0};
    if (xc[2] < zmin2 || xc[2] > zmin2) {
        damps[2] = xc[2] < zmin2 ? std::fabs(xc[2] - zmin2) : 0.0;
        if (xc[0] > xmax2 && xc[1] > ymax2) { damps[0] = std::fabs(xc[0] - xmax2); damps[1] = std::fabs(xc[1] - ymax2); }
        else if (xc[0] > xmax2 && xc[1] < ymin2) { damps[0] = std::fabs(xc[0] - xmax2); damps[1] = std::fabs(xc[1] - ymin2); }
        else if (xc[0] < xmin2 && xc[1] < ymin2) { damps[0] = std::fabs(xc[0] - xmin2); damps[1] = std::fabs(xc[1] - ymin2); }
        else if (xc[0] < xmin2 && xc[1] > xmax2) { damps[0] = std::fabs(xc[0] - xmin2); damps[1] = std::fabs(xc[1] - ymax2); }
        else if (xc[0] > xmax2 && xc[1] > ymin2 && xc[1] < ymax2) { damps[0] = std::fabs(xc[0] - xmax2); damps[1] = 0.0; }
        else if (xc[1] < ymin2 && xc[0] > xmin2 && xc[0] < xmax2) { damps[0] = 0.0; damps[1] = std::fabs(xc[1] - ymin2); }
        else if (xc[0] < xmin2 && xc[1] > ymin2 && xc[1] < ymax2) { damps[0] = std::fabs(xc[0] - xmin2); damps[1] = 0.0; }
        else if (xc[1] > ymax2 && xc[0] > xmin2 && xc[0] < xmax2) { damps[0] = 0.0; damps[1] = std::fabs(xc[1] - ymax2); }
    }
    for (int i = 0; i < 3; i++) {
        const double delta = (double)s.nPML * s.PMLb[5 + i];
        d[i] = 3 * s.vmaxPML / 2 / delta * std::log(1 / s.R) * std::pow(damps[i] / delta, (double)2.0f);
    }
}

// qconstant, src/qconstant.f90:3-34: the coefficient literals are default
// REAL there, so each is the float value; x**(-0.5) is 1/sqrt(x).
static void qconstant(double Q, int k, double pi, double *rtaok, double *rwk, double *c1)
{
    static const float taok[8] = {1.72333e-3f, 1.80701e-3f, 5.38887e-3f, 1.99322e-2f,
                                  8.49833e-2f, 4.09335e-1f, 2.05951f, 13.2629f};
    static const float alfk[8] = {1.66958e-2f, 3.81644e-2f, 9.84666e-3f, -1.36803e-2f,
                                  -2.85125e-2f, -5.37309e-2f, -6.65035e-2f, -1.33696e-1f};
    static const float betk[8] = {8.98758e-2f, 6.84635e-2f, 9.67052e-2f, 1.20172e-1f,
                                  1.30728e-1f, 1.38746e-1f, 1.40705e-1f, 2.14647e-1f};
    double kapa = (double)3.071f + (double)1.433f * std::pow(Q, (double)(-1.158f)) * std::log(Q / 5.0);
    kapa = kapa / (1.0 + (double)0.415f * Q);
    *rwk = kapa * (kapa * (double)alfk[k - 1] + (double)betk[k - 1]);
    *rtaok = (double)taok[k - 1];
    const double ref = 2.0 * pi;
    const double ak0 = 1.0 - *rwk * 8.0 / (1.0 + (*rtaok * ref) * (*rtaok * ref));
    const double bk0 = *rwk * 8.0 * ref * *rtaok / (1.0 + (*rtaok * ref) * (*rtaok * ref));
    const double rs = 1.0 / std::sqrt(ak0 * ak0 + bk0 * bk0);
    *c1 = (0.5 * rs) * (1.0 + ak0 * rs);
}

// During eqd_commit arrays come from the caller; afterwards from the kept copies.
static const double *F64(eqd_handle *h, int f)
{
    auto k = h->fd.find(f);
    if (k != h->fd.end()) return k->second.data();
    auto it = h->src.find(f);
    return it == h->src.end() ? nullptr : (const double *)it->second;
}
static const int32_t *I32(eqd_handle *h, int f)
{
    auto k = h->fi.find(f);
    if (k != h->fi.end()) return k->second.data();
    auto it = h->src.find(f);
    return it == h->src.end() ? nullptr : (const int32_t *)it->second;
}
// Groups rows whose values are bit-identical: idx[t] is the row's class and
// table holds one copy per class.  Gives up (returns false) when there are
// more than max_classes classes, in which case per-element arrays are used.
template <typename R, typename RowFn>
static bool classify_rows(int64_t n, int width, RowFn row, int64_t max_classes, std::vector<uint32_t> &idx,
                          std::vector<R> &table)
{
    idx.assign(n, 0);
    table.clear();
    std::unordered_map<uint64_t, std::vector<uint32_t>> buckets;
    std::vector<R> tmp(width);
    int64_t nclass = 0;
    for (int64_t t = 0; t < n; t++) {
        row(t, tmp.data());
        uint64_t hsh = 1469598103934665603ull;
        const unsigned char *b = (const unsigned char *)tmp.data();
        for (size_t i = 0; i < sizeof(R) * width; i++) hsh = (hsh ^ b[i]) * 1099511628211ull;
        auto &bk = buckets[hsh];
        int64_t c = -1;
        for (uint32_t k : bk)
            if (std::memcmp(table.data() + (size_t)k * width, tmp.data(), sizeof(R) * width) == 0) {
                c = k;
                break;
            }
        if (c < 0) {
            if (nclass >= max_classes) return false;
            c = nclass++;
            bk.push_back((uint32_t)c);
            table.insert(table.end(), tmp.begin(), tmp.end());
        }
        idx[t] = (uint32_t)c;
    }
    return true;
}

// Compact hourglass operators of one element (phi_compact): for each mode
// the least-squares fit of (1/s_m, h_m/s_m) to the element's eight Phi values
// given dN/dx, in long double on column-scaled normal equations.  Returns
// false when the rebuilt values miss the given ones by more than 1e-11 of
// max(1, |Phi|): Phi that is not of the reference's form keeps its table.
static bool fit_phi_compact(const double shp[24], const double phi[32], double c[16])
{
    for (int m = 0; m < 4; m++) {
        long double A[8][4], b[8];
        for (int a = 0; a < 8; a++) {
            A[a][0] = ((kGammaNeg[m] >> a) & 1u) ? -1.0L : 1.0L;
            for (int k = 0; k < 3; k++) A[a][1 + k] = -(long double)shp[a * 3 + k];
            b[a] = phi[m * 8 + a];
        }
        long double sc[4];
        for (int j = 0; j < 4; j++) {
            long double q = 0;
            for (int a = 0; a < 8; a++) q += A[a][j] * A[a][j];
            sc[j] = q > 0 ? 1.0L / std::sqrt(q) : 1.0L;
            for (int a = 0; a < 8; a++) A[a][j] *= sc[j];
        }
        long double N[4][5];
        for (int i = 0; i < 4; i++) {
            for (int j = 0; j < 4; j++) {
                long double q = 0;
                for (int a = 0; a < 8; a++) q += A[a][i] * A[a][j];
                N[i][j] = q;
            }
            long double q = 0;
            for (int a = 0; a < 8; a++) q += A[a][i] * b[a];
            N[i][4] = q;
        }
        for (int i = 0; i < 4; i++) {
            int piv = i;
            for (int r = i + 1; r < 4; r++)
                if (std::fabs(N[r][i]) > std::fabs(N[piv][i])) piv = r;
            if (!(std::fabs(N[piv][i]) > 1e-30L)) return false;
            if (piv != i)
                for (int j = 0; j < 5; j++) std::swap(N[i][j], N[piv][j]);
            for (int r = i + 1; r < 4; r++) {
                const long double f = N[r][i] / N[i][i];
                for (int j = i; j < 5; j++) N[r][j] -= f * N[i][j];
            }
        }
        long double x[4];
        for (int i = 3; i >= 0; i--) {
            long double q = N[i][4];
            for (int j = i + 1; j < 4; j++) q -= N[i][j] * x[j];
            x[i] = q / N[i][i];
        }
        c[m] = (double)(x[0] * sc[0]);
        for (int k = 0; k < 3; k++) c[4 + 3 * m + k] = (double)(x[1 + k] * sc[1 + k]);
    }
    double big = 1.0;
    for (int i = 0; i < 32; i++) big = std::max(big, std::fabs(phi[i]));
    for (int m = 0; m < 4; m++)
        for (int a = 0; a < 8; a++) {
            const double r = phi_compact<double>(c, m, a, shp[a * 3], shp[a * 3 + 1], shp[a * 3 + 2]);
            if (!(std::fabs(r - phi[m * 8 + a]) <= 1e-11 * big)) return false;
        }
    return true;
}

// The element operators of one class (regular or PML) on the device: either
// lossless classes (both the operator rows -- kOpW values -- and the
// hourglass Phi rows -- 32 values -- collapse to at most n/16 distinct rows)
// or, when they do not, one AoSoA array of operator rows per element in
// blocks of 64 elements: value k of element t at ((t / 64) * W + k) * 64 + t % 64.
// When every element's Phi has the reference's form (fit_phi_compact), Phi is
// carried in compact form: per-element rows of W = kOpBc values end with the
// 16 compact values, and the class tables hold Phi rebuilt from them in the
// kernels' precision and order (phi_compact), so both layouts compute the same
// numbers.  Otherwise (or with EQD_FULL_PHI=1) rows of W = kOpB values carry
// the given Phi.  shp and phi are the class's [24][n] and [32][n] arrays.
// The geometry row of one hexahedron (kOpG values: the coordinates of local
// nodes 1..7 relative to node 0, lambda, mu, Vp, rho), if its given
// operators are those of its coordinates: the one-time precompute
// (eqd_precompute.h) repeated here from xl must give dN/dx, det, SS and Phi
// within 1e-12 of each quantity's size.
template <typename R>
static bool geometry_row(const double xl[3][8], double lam, double mu, double vp, double rho, double w,
                         const double *shp, double det, const double *ss, const double *phi, R row[kOpG])
{
    double sh[24], dt, em[24], s6[6], ph[32];
    if (!eqd::elem_precompute(xl, false, rho, lam, mu, w, sh, &dt, em, s6, ph)) return false;
    auto close = [](const double *a, const double *b, int k) {
        double m = 0, d = 0;
        for (int i = 0; i < k; i++) {
            m = std::max(m, std::fabs(b[i]));
            d = std::max(d, std::fabs(a[i] - b[i]));
        }
        return d <= 1e-12 * m;
    };
    if (!close(sh, shp, 24) || !close(&dt, &det, 1) || !close(s6, ss, 6) || !close(ph, phi, 32)) return false;
    for (int a = 1; a < 8; a++)
        for (int j = 0; j < 3; j++) row[(a - 1) * 3 + j] = (R)(xl[j][a] - xl[j][0]);
    row[kGeoMat] = (R)lam;
    row[kGeoMat + 1] = (R)mu;
    row[kGeoMat + 2] = (R)vp;
    row[kGeoMat + 3] = (R)rho;
    return true;
}

//
// Geometry rows (FP32 bulk mode, allow_geom): where neither class tables nor
// anything shorter applies, and geomrow(t, row) confirms that the element's
// given operators are those of its corner coordinates and fills its row
// (kOpG values), the kernels rebuild the operators from the coordinates
// (kOpsGeom): 25 values per element instead of 51.
template <typename R, typename E, typename OpRow, typename GeomRow>
static int upload_operators(eqd_handle *h, E &X, int64_t n, bool no_classes, OpRow oprow,
                            const std::vector<double> &shp, const std::vector<double> &phi, int64_t *n_op_classes,
                            int64_t *n_phi_classes, int32_t *compact, bool allow_geom, GeomRow geomrow)
{
    std::vector<R> cmp;
    int ok = 1;
    const char *fe = std::getenv("EQD_FULL_PHI");
    if (!(fe && fe[0] == '1') && n > 0) {
        cmp.resize((size_t)n * 16);
#pragma omp parallel for schedule(static) reduction(min : ok)
        for (int64_t t = 0; t < n; t++) {
            double sh[24], ph[32], c[16];
            for (int i = 0; i < 24; i++) sh[i] = shp[i * n + t];
            for (int i = 0; i < 32; i++) ph[i] = phi[i * n + t];
            if (!fit_phi_compact(sh, ph, c)) ok = 0;
            for (int i = 0; i < 16; i++) cmp[(size_t)t * 16 + i] = (R)c[i];
        }
    } else {
        ok = 0;
    }
    *compact = ok;
    auto phirow = [&](int64_t t, R *row) {
        if (ok) {
            for (int m = 0; m < 4; m++)
                for (int a = 0; a < 8; a++)
                    row[m * 8 + a] = phi_compact<R>(&cmp[(size_t)t * 16], m, a, (R)shp[(a * 3) * n + t],
                                                    (R)shp[(a * 3 + 1) * n + t], (R)shp[(a * 3 + 2) * n + t]);
        } else {
            for (int i = 0; i < 32; i++) row[i] = (R)phi[i * n + t];
        }
    };
    std::vector<uint32_t> ocls, pcls;
    std::vector<R> otab, ptab;
    const int64_t maxc = no_classes ? 0 : std::max<int64_t>(1, n / 16);
    const bool ok_cls = classify_rows<R>(n, kOpW, oprow, maxc, ocls, otab) &&
                        classify_rows<R>(n, 32, phirow, maxc, pcls, ptab);
    int rc;
    X.opb_compact = 0;
    if (ok_cls) {
        uint32_t *pc;
        R *pt;
        if ((rc = dupload(h, &pc, ocls)) || (rc = dupload(h, &pt, otab))) return rc;
        X.opc = pc;
        X.optab = pt;
        X.n_opc = (int32_t)(otab.size() / kOpW);
        if ((rc = dupload(h, &pc, pcls)) || (rc = dupload(h, &pt, ptab))) return rc;
        X.phic = pc;
        X.phitab = pt;
        X.n_phic = (int32_t)(ptab.size() / 32);
        *n_op_classes = X.n_opc;
        *n_phi_classes = X.n_phic;
        return 0;
    }
    const int64_t nb = (n + 63) / 64;
    X.opb_geom = 0;
    if (allow_geom && ok) {
        std::vector<R> g((size_t)(nb * 64 * kOpG), (R)0);
        int gok = 1;
#pragma omp parallel for schedule(static) reduction(min : gok)
        for (int64_t t = 0; t < n; t++) {
            R row[kOpG];
            if (!geomrow(t, row)) {
                gok = 0;
                continue;
            }
            R *dst = g.data() + (t / 64) * kOpG * 64 + (t % 64);
            for (int k = 0; k < kOpG; k++) dst[k * 64] = row[k];
        }
        if (gok) {
            R *pb;
            if ((rc = dupload(h, &pb, g))) return rc;
            X.opb = pb;
            X.opb_compact = 1;
            X.opb_geom = 1;
            X.optab = nullptr;
            X.phitab = nullptr;
            *n_op_classes = 0;
            *n_phi_classes = 0;
            return 0;
        }
    }
    const int W = ok ? kOpBc : kOpB;
    std::vector<R> b((size_t)(nb * 64 * W), (R)0);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n; t++) {
        R row[kOpB];
        oprow(t, row);
        if (ok) {
            for (int i = 0; i < 16; i++) row[kOpW + i] = cmp[(size_t)t * 16 + i];
        } else {
            phirow(t, row + kOpW);
        }
        R *dst = b.data() + (t / 64) * W * 64 + (t % 64);
        for (int k = 0; k < W; k++) dst[k * 64] = row[k];
    }
    R *pb;
    if ((rc = dupload(h, &pb, b))) return rc;
    X.opb = pb;
    X.opb_compact = ok;
    X.optab = nullptr;
    X.phitab = nullptr;
    *n_op_classes = 0;
    *n_phi_classes = 0;
    return 0;
}

// Element order within a class: bricks of 4 (x) x 8 (z) x 32 (y) elements,
// y fastest inside a brick and bricks in y, then z, then x order, so that the
// elements that share a node run close together in time and the node's data
// is still in L2 when its next element needs it.  Grid indices come from the
// element centroids: x and z as ranks of the distinct centroid coordinates
// (to the millimetre), y as the rank within each (x, z) column, which holds
// for columns shifted by a rough fault surface.  Results do not depend on the
// order: every node sums its elements in ascending reference id.  Selected by
// EQD_ELEMENT_ORDER=1; measured no faster than the reference order on the
// 50 m and rough-fault boxes (the element kernels' L2 misses are mostly their
// compulsory streams), so the reference order is the default.
static void brick_order(std::vector<int32_t> &ids, const double *coor, const int32_t *ien)
{
    const int64_t n = (int64_t)ids.size();
    if (n == 0) return;
    std::vector<int64_t> qx(n), qz(n);
    std::vector<double> cy(n);
    for (int64_t t = 0; t < n; t++) {
        const int64_t e = ids[t];
        double c[3] = {0.0, 0.0, 0.0};
        for (int a = 0; a < 8; a++)
            for (int k = 0; k < 3; k++) c[k] += coor[(int64_t)(ien[e * 8 + a] - 1) * 3 + k];
        qx[t] = std::llround(c[0] / 8.0 * 1000.0);
        qz[t] = std::llround(c[2] / 8.0 * 1000.0);
        cy[t] = c[1] / 8.0;
    }
    auto ranks = [n](const std::vector<int64_t> &q) {
        std::vector<int64_t> u(q);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        std::vector<int32_t> r(n);
        for (int64_t t = 0; t < n; t++) r[t] = (int32_t)(std::lower_bound(u.begin(), u.end(), q[t]) - u.begin());
        return r;
    };
    const std::vector<int32_t> xi = ranks(qx), zi = ranks(qz);
    std::vector<int64_t> ord(n);
    for (int64_t t = 0; t < n; t++) ord[t] = t;
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
        if (xi[a] != xi[b]) return xi[a] < xi[b];
        if (zi[a] != zi[b]) return zi[a] < zi[b];
        if (cy[a] != cy[b]) return cy[a] < cy[b];
        return ids[a] < ids[b];
    });
    std::vector<int32_t> yi(n);
    for (int64_t k = 0; k < n; k++) {
        const int64_t t = ord[k];
        const bool same = k > 0 && xi[ord[k - 1]] == xi[t] && zi[ord[k - 1]] == zi[t];
        yi[t] = same ? yi[ord[k - 1]] + 1 : 0;
    }
    auto key = [&](int64_t t) {
        return std::make_tuple(xi[t] / 4, zi[t] / 8, yi[t] / 32, xi[t] % 4, zi[t] % 8, yi[t] % 32, ids[t]);
    };
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return key(a) < key(b); });
    std::vector<int32_t> out(n);
    for (int64_t k = 0; k < n; k++) out[k] = ids[ord[k]];
    ids.swap(out);
}

template <typename T>
static void keep_copy(eqd_handle *h, int f, std::map<int, std::vector<T>> &dst)
{
    auto it = h->src.find(f);
    if (it == h->src.end()) return;
    FieldSpec sp;
    field_spec(h->s, f, sp);
    const T *p = (const T *)it->second;
    dst[f].assign(p, p + sp.count);
}

namespace eqd {

int settle(eqd_handle *h)
{
    if (!h || h->inflight == 0) return EQD_OK;
    const int32_t n = h->inflight;
    h->inflight = 0;
    const int rc = engine_finish(h->eng, n);
    if (rc) h->err = h->eng->err;
    return rc;
}

}  // namespace eqd

extern "C" {

int eqd_create(const eqd_scalars *s, eqd_handle **out)
{
    if (!s || !out) return EQD_EINVAL;
    *out = nullptr;
    if (s->nn <= 0 || s->ne <= 0 || s->neq <= 0 || s->nf < 0 || s->non < 0 || s->noff < 0 || s->nstep <= 0)
        return EQD_EINVAL;
    if (s->precision != EQD_PREC_FP64 && s->precision != EQD_PREC_FP32) return EQD_EINVAL;
    eqd_handle *h = new eqd_handle();
    h->s = *s;
    *out = h;
    return EQD_OK;
}

int eqd_set(eqd_handle *h, int32_t field, const void *host, int64_t count)
{
    if (!h) return EQD_EINVAL;
    SETTLE(h);
    FieldSpec sp;
    if (!field_spec(h->s, field, sp)) return fail(h, EQD_EINVAL, "unknown field %d", field);
    if (count != sp.count)
        return fail(h, EQD_EINVAL, "field %d: count %lld, expected %lld", field, (long long)count, (long long)sp.count);
    if (h->committed) return fail(h, EQD_ESTATE, "field %d: arrays are set before eqd_commit", field);
    if (count > 0 && !host) return fail(h, EQD_EINVAL, "field %d: null pointer", field);
    h->src[field] = host;
    return EQD_OK;
}



int eqd_commit(eqd_handle *h)
{
    if (!h) return EQD_EINVAL;
    if (h->committed) return fail(h, EQD_ESTATE, "already committed");
    const eqd_scalars &s = h->s;
    const int required[] = {EQD_MESHCOOR, EQD_EQNUMINDEXARR, EQD_EQNUMSTARTINDEXLOC, EQD_NUMOFDOFPERNODEARR,
                            EQD_NODEELEMIDRELATION, EQD_ELEMTYPEARR, EQD_MAT, EQD_ELESHP, EQD_ELEDET, EQD_ELEMASS,
                            EQD_SS, EQD_PHI, EQD_STRESSCOMPINDEXARR, EQD_ELEPOREP, EQD_NODALMASSARR, EQD_FNMS,
                            EQD_NSMP, EQD_UN, EQD_US, EQD_UD, EQD_ARN, EQD_STRESSARR, EQD_PSTRAIN, EQD_V1,
                            EQD_FRIC, EQD_FNFT};
    for (int f : required) {
        FieldSpec sp;
        field_spec(s, f, sp);
        if (!h->src.count(f) && sp.count > 0) return fail(h, EQD_EINVAL, "field %d was not set", f);
    }
    if (s.non > 0 && !h->src.count(EQD_ANONFS)) return fail(h, EQD_EINVAL, "anonfs was not set");
    if (s.noff > 0 && !h->src.count(EQD_IDHIST)) return fail(h, EQD_EINVAL, "idhist was not set");
    if (s.C_Q == 1 && s.C_elastic == 0) return fail(h, EQD_EINVAL, "Q model can only work with elastic code");
    if (s.friclaw < 1 || s.friclaw > 5) return fail(h, EQD_EINVAL, "friclaw %d is not supported (1-5)", s.friclaw);
    // Split-node pairs, checked before anything touches the device.  Every
    // pair updates its two nodes in one thread, all pairs at once (the
    // reference goes through them one after another), so a node may belong to
    // one pair only and a pair needs two distinct nodes.
    {
        const int32_t *nsmp = I32(h, EQD_NSMP), *ndof = I32(h, EQD_NUMOFDOFPERNODEARR);
        std::vector<uint8_t> seen(s.nf > 0 ? s.nn : 0, 0);
        for (int64_t i = 0; i < s.nf; i++) {
            if (nsmp[i * 2] == nsmp[i * 2 + 1])
                return fail(h, EQD_EINVAL, "pair %lld: slave and master are the same node %d", (long long)i + 1,
                            nsmp[i * 2]);
            for (int j = 0; j < 2; j++) {
                const int32_t nd = nsmp[i * 2 + j];
                if (nd < 1 || nd > s.nn)
                    return fail(h, EQD_EINVAL, "nsmp(%d,%lld) = %d out of range", j + 1, (long long)i + 1, nd);
                if (ndof[nd - 1] != 3) return fail(h, EQD_EINVAL, "split node %d is not a 3-component node", nd);
                if (seen[nd - 1])
                    return fail(h, EQD_EINVAL, "node %d belongs to more than one split-node pair (pair %lld)", nd,
                                (long long)i + 1);
                seen[nd - 1] = 1;
            }
        }
    }
    HIPCHK(h, hipSetDevice(s.device));
    // the handle's own engine (streams, launch tables); built at the first run
    if (!h->eng) {
        h->eng = new eqd_engine();
        int rc = engine_init(h->eng, {h});
        if (rc) {
            h->err = h->eng->err;
            return rc;
        }
        h->stream = h->eng->st;
        h->eng->timing = h->timing;
    }

    const int64_t nn = s.nn, ne = s.ne;
    const double *coor = F64(h, EQD_MESHCOOR);
    const int32_t *eqi = I32(h, EQD_EQNUMINDEXARR), *esl = I32(h, EQD_EQNUMSTARTINDEXLOC),
                  *ndof = I32(h, EQD_NUMOFDOFPERNODEARR), *ien = I32(h, EQD_NODEELEMIDRELATION),
                  *etype = I32(h, EQD_ELEMTYPEARR), *sci = I32(h, EQD_STRESSCOMPINDEXARR), *nsmp = I32(h, EQD_NSMP);
    const double *mat = F64(h, EQD_MAT), *shp = F64(h, EQD_ELESHP), *det = F64(h, EQD_ELEDET),
                 *emass = F64(h, EQD_ELEMASS), *ss = F64(h, EQD_SS), *phi = F64(h, EQD_PHI),
                 *porep = F64(h, EQD_ELEPOREP), *nmass = F64(h, EQD_NODALMASSARR), *fnms = F64(h, EQD_FNMS),
                 *stress = F64(h, EQD_STRESSARR), *pstrain = F64(h, EQD_PSTRAIN), *v1 = F64(h, EQD_V1);
    const double *vel0 = F64(h, EQD_VELARR), *disp0 = F64(h, EQD_DISPARR), *force0 = F64(h, EQD_NODALFORCEARR);

    auto eq_of = [&](int64_t node, int j) -> int32_t {  // node 0-based, j 1-based
        const int64_t idx = (int64_t)esl[node] + j;     // 1-based position in eqNumIndexArr
        if (idx < 1 || idx > s.size_eqidx) return -2;
        return eqi[idx - 1];
    };

    // ---------------- nodes
    std::vector<double> mass(nn, 0.0);
    std::vector<uint8_t> is_split(nn, 0);
    for (int64_t i = 0; i < s.nf; i++)
        for (int j = 0; j < 2; j++) is_split[nsmp[i * 2 + j] - 1] = 1;
    h->pml_node_index.assign(nn, -1);
    std::vector<double> vel(3 * nn, 0.0), disp(3 * nn, 0.0), acc(3 * nn, 0.0);
    std::vector<double> pdamp, pv12, pacc;
    for (int64_t n = 0; n < nn; n++) {
        for (int k = 0; k < 3; k++) {
            vel[k * nn + n] = vel0 ? vel0[n * 3 + k] : 0.0;
            disp[k * nn + n] = disp0 ? disp0[n * 3 + k] : 0.0;
        }
        if (ndof[n] == 3) {
            for (int j = 1; j <= 3; j++) {
                const int32_t e = eq_of(n, j);
                if (e < 1 || e > s.neq)
                    return fail(h, EQD_EINVAL, "3-component node %lld has equation %d (fixed 3-component nodes are not supported)",
                                (long long)n + 1, e);
                if (j == 1) mass[n] = nmass[e - 1];
                else if (nmass[e - 1] != mass[n])
                    return fail(h, EQD_EINVAL, "node %lld: unequal component masses", (long long)n + 1);
                vel[(j - 1) * nn + n] = v1[e - 1];
                acc[(j - 1) * nn + n] = force0 ? force0[e - 1] : 0.0;
            }
            h->node3_all.push_back((int32_t)n);
            if (!is_split[n]) h->node3_free.push_back((int32_t)n);
        } else if (ndof[n] == 12) {
            const int32_t e1 = eq_of(n, 1);
            if (e1 == -1) {
                for (int j = 2; j <= 12; j++)
                    if (eq_of(n, j) != -1) return fail(h, EQD_EINVAL, "node %lld: partly fixed PML node", (long long)n + 1);
                continue;  // fixed boundary node: velocity and displacement stay as given
            }
            const int32_t q = (int32_t)h->pml_node_ids.size();
            h->pml_node_index[n] = q;
            h->pml_node_ids.push_back((int32_t)n);
            for (int j = 1; j <= 12; j++) {
                const int32_t e = eq_of(n, j);
                if (e < 1 || e > s.neq) return fail(h, EQD_EINVAL, "node %lld: partly fixed PML node", (long long)n + 1);
                if (j == 1) mass[n] = nmass[e - 1];
                else if (nmass[e - 1] != mass[n])
                    return fail(h, EQD_EINVAL, "node %lld: unequal component masses", (long long)n + 1);
                pv12.push_back(v1[e - 1]);
                pacc.push_back(force0 ? force0[e - 1] : 0.0);
            }
            double d[3];
            pml_node_damping(s, coor[n * 3], coor[n * 3 + 1], coor[n * 3 + 2], d);
            for (int k = 0; k < 3; k++) {
                if (d[k] < 0.0) return fail(h, EQD_EINVAL, "wrong dv");
                pdamp.push_back(d[k]);
            }
        } else {
            return fail(h, EQD_EINVAL, "node %lld has %d components", (long long)n + 1, ndof[n]);
        }
    }
    const int64_t npn = (int64_t)h->pml_node_ids.size();
    // per-PML-node arrays to component-major (then to blocks at upload)
    auto tr = [](const std::vector<double> &a, int64_t cnt, int w) {
        std::vector<double> b(a.size());
        for (int64_t q = 0; q < cnt; q++)
            for (int k = 0; k < w; k++) b[k * cnt + q] = a[q * w + k];
        return b;
    };
    pv12 = tr(pv12, npn, 12);
    pacc = tr(pacc, npn, 12);
    pdamp = tr(pdamp, npn, 3);

    // ---------------- elements
    const int nst = s.C_Q == 1 ? 12 : 6;
    h->nreg_stress = nst;
    for (int64_t e = 0; e < ne; e++) {
        const int t = etype[e];
        if (t == 1 || t > 10) h->reg_ids.push_back((int32_t)e);
        else if (t == 2) h->pml_ids.push_back((int32_t)e);
        else return fail(h, EQD_EINVAL, "element %lld has type %d", (long long)e + 1, t);
        for (int a = 0; a < 8; a++) {
            const int32_t nd = ien[e * 8 + a];
            if (nd < 1 || nd > nn) return fail(h, EQD_EINVAL, "element %lld node %d out of range", (long long)e + 1, nd);
            if ((t == 1 || t > 10) && ndof[nd - 1] != 3)
                return fail(h, EQD_EINVAL, "regular element %lld touches a PML node", (long long)e + 1);
        }
        const int need = t == 2 ? 21 : 12;
        if (sci[e] < 0 || (int64_t)sci[e] + need > s.size_stress + (t == 2 ? 0 : 6))
            return fail(h, EQD_EINVAL, "element %lld: stress index out of range", (long long)e + 1);
    }
    {
        const char *o = std::getenv("EQD_ELEMENT_ORDER");
        if (o && o[0] == '1') {
            brick_order(h->reg_ids, coor, ien);
            brick_order(h->pml_ids, coor, ien);
        }
    }
    const int64_t nr = (int64_t)h->reg_ids.size(), np = (int64_t)h->pml_ids.size();
    if (nr > (int64_t)kEntIdxMask || np > (int64_t)kEntIdxMask) return fail(h, EQD_EINVAL, "too many elements in one class");
    h->n_reg = (int32_t)nr;
    h->n_pml = (int32_t)np;
    const bool fp32 = s.precision == EQD_PREC_FP32;
    const char *nc_env = std::getenv("EQD_NO_OPERATOR_CLASSES");
    const bool no_classes = nc_env && nc_env[0] == '1';
    // geometry rows: FP32 bulk mode, hexahedra only (EQD_GEOM_OPS=0 keeps the
    // stored operators, 2 rebuilds the regular elements' only)
    const char *go_env = std::getenv("EQD_GEOM_OPS");
    const bool geom_ok = fp32 && !(go_env && go_env[0] == '0');
    const bool geom_pml = geom_ok && !(go_env && go_env[0] == '2');
    bool reg_wedges = false;
    for (int32_t e : h->reg_ids)
        if (etype[e] == 11 || etype[e] == 12) reg_wedges = true;

    // regular class SoA
    {
        std::vector<int32_t> r_ien(8 * nr);
        std::vector<double> r_shp(24 * nr), r_det(nr), r_lam(nr), r_mu(nr), r_vp(nr), r_rho(nr), r_st(nst * nr),
            r_ss(6 * nr), r_phi(32 * nr), r_em(24 * nr), r_porep(nr), r_pstrain(nr);
        std::vector<uint8_t> r_q(s.C_Q == 1 ? nr : 0);
        int qbad = 0;
#pragma omp parallel for schedule(static) reduction(max : qbad)
        for (int64_t t = 0; t < nr; t++) {
            const int64_t e = h->reg_ids[t];
            for (int a = 0; a < 8; a++) r_ien[a * nr + t] = ien[e * 8 + a] - 1;
            for (int i = 0; i < 24; i++) {
                r_shp[i * nr + t] = shp[e * 24 + i];
                r_em[i * nr + t] = emass[e * 24 + i];
            }
            r_det[t] = det[e];
            r_vp[t] = mat[0 * ne + e];
            r_rho[t] = mat[2 * ne + e];
            r_lam[t] = mat[3 * ne + e];
            r_mu[t] = mat[4 * ne + e];
            for (int i = 0; i < 6; i++) r_ss[i * nr + t] = ss[e * 6 + i];
            for (int m = 0; m < 4; m++)
                for (int a = 0; a < 8; a++) r_phi[(m * 8 + a) * nr + t] = phi[e * 32 + m * 8 + a];
            for (int i = 0; i < nst; i++) {
                const int64_t idx = (int64_t)sci[e] + i;
                r_st[i * nr + t] = idx < s.size_stress ? stress[idx] : 0.0;
            }
            r_porep[t] = porep[e];
            r_pstrain[t] = pstrain[e];
            if (s.C_Q == 1) {
                // Q law and relaxation mechanism of src/calcElemKU.f90:78-103
                double xc[3] = {0, 0, 0};
                for (int i = 0; i < 3; i++)
                    for (int a = 0; a < 8; a++) xc[i] = xc[i] + coor[(int64_t)(ien[e * 8 + a] - 1) * 3 + i];
                for (int i = 0; i < 3; i++) xc[i] = xc[i] / 8.0;
                const int qsel = xc[2] > -1000.0 ? 0 : 1;
                const int ip = (int)((xc[0] - (s.PMLb[1] + s.dx / 2)) / s.dx + 1);
                const int iq = (int)((xc[1] - (s.PMLb[3] + s.dx / 2)) / s.dx + 1);
                const int ir = (int)((xc[2] - (s.PMLb[4] + s.dx / 2)) / s.dx + 1);
                const int k = 1 + ip % 2 + 2 * (iq % 2) + 4 * (ir % 2);
                if (k < 1 || k > 8) {
                    qbad = 1;
                    continue;
                }
                r_q[t] = (uint8_t)(qsel * 8 + (k - 1));
            }
        }
        if (qbad) return fail(h, EQD_EINVAL, "Q relaxation mechanism outside 1..8 (element centre outside the PML box?)");
        std::vector<double> qtab(16 * 6, 0.0);
        if (s.C_Q == 1) {
            for (int qsel = 0; qsel < 2; qsel++)
                for (int k = 1; k <= 8; k++) {
                    const double Qs = qsel == 0 ? 10.0 : 50.0, Qp = qsel == 0 ? 20.0 : 100.0;
                    double taok, wkp, wks, cv, cs;
                    qconstant(Qp, k, s.pi, &taok, &wkp, &cv);
                    qconstant(Qs, k, s.pi, &taok, &wks, &cs);
                    double *row = &qtab[(qsel * 8 + k - 1) * 6];
                    row[0] = std::exp(-s.dt / taok);
                    row[1] = 1 - std::exp(-s.dt / taok);
                    row[2] = wkp * 8.0;
                    row[3] = wks * 8.0;
                    row[4] = cv;
                    row[5] = cs;
                }
        }
        int rc;
        int32_t *d_ien;
        if ((rc = dupload(h, &d_ien, to_blocks<int32_t>(r_ien, nr, 8)))) return rc;
        uint8_t *d_q = nullptr;
        double *d_qtab = nullptr, *d_porep, *d_pstrain;
        if ((rc = dupload(h, &d_q, r_q))) return rc;
        if ((rc = dupload(h, &d_qtab, qtab))) return rc;
        if ((rc = dupload(h, &d_porep, r_porep))) return rc;
        if ((rc = dupload(h, &d_pstrain, r_pstrain))) return rc;
        auto fill = [&](auto &E) -> int {
            using R = std::remove_pointer_t<decltype(E.ef)>;
            auto up = [&](const R **dst, const std::vector<double> &src, int w) -> int {
                R *p;
                int r2 = dupload(h, &p, to_blocks<R>(src, nr, w));
                *dst = p;
                return r2;
            };
            int r2;
            E.n = nr;
            E.ien = d_ien;
            // operators: lossless classes, or per element in AoSoA blocks
            // (EQD_NO_OPERATOR_CLASSES=1 forces the latter; used to test that both give equal results)
            if ((r2 = upload_operators<R>(h, E, nr, no_classes, [&](int64_t t, R *row) {
                    for (int i = 0; i < 24; i++) row[kOpShp + i] = (R)r_shp[i * nr + t];
                    row[kOpDet] = (R)r_det[t];
                    row[kOpLam] = (R)r_lam[t];
                    row[kOpMu] = (R)r_mu[t];
                    row[kOpVp] = (R)r_vp[t];
                    row[kOpRho] = (R)r_rho[t];
                    for (int i = 0; i < 6; i++) row[kOpSs + i] = (R)r_ss[i * nr + t];
                }, r_shp, r_phi, &h->op_classes_reg, &h->phi_classes_reg, &h->phi_compact_reg,
                geom_ok && !reg_wedges, [&](int64_t t, R *row) {
                    const int64_t e = h->reg_ids[t];
                    double xl[3][8], sh[24], s6[6], ph[32];
                    for (int a = 0; a < 8; a++)
                        for (int j = 0; j < 3; j++) xl[j][a] = coor[(int64_t)(ien[e * 8 + a] - 1) * 3 + j];
                    for (int i = 0; i < 24; i++) sh[i] = r_shp[i * nr + t];
                    for (int i = 0; i < 6; i++) s6[i] = r_ss[i * nr + t];
                    for (int i = 0; i < 32; i++) ph[i] = r_phi[i * nr + t];
                    return geometry_row<R>(xl, r_lam[t], r_mu[t], r_vp[t], r_rho[t], s.w, sh, r_det[t], s6, ph, row);
                })))
                return r2;
            if (s.rdampm != 0.0 || s.C_elastic == 0) {
                if ((r2 = up(&E.emass, r_em, 24))) return r2;
            }
            R *p;
            if ((r2 = dupload(h, &p, to_blocks<R>(r_st, nr, nst)))) return r2;
            E.stress = p;
            if ((r2 = dalloc(h, &p, (nr + 63) / 64 * 64 * kEfReg))) return r2;
            E.ef = p;
            E.porep = d_porep;
            E.pstrain = d_pstrain;
            E.qidx = d_q;
            E.qtab = d_qtab;
            return 0;
        };
        if (fp32) {
            if ((rc = fill(h->E32))) return rc;
        } else {
            if ((rc = fill(h->E64))) return rc;
        }
    }
    // PML class SoA
    {
        std::vector<int32_t> p_ien(8 * np);
        std::vector<double> p_shp(24 * np), p_det(np), p_lam(np), p_mu(np), p_vp(np), p_rho(np), p_st(21 * np),
            p_ss(6 * np), p_phi(32 * np), p_em(24 * np), p_damp(3 * np);
        std::vector<uint8_t> p_mask(np);
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < np; p++) {
            const int64_t e = h->pml_ids[p];
            uint8_t m = 0;
            double xc[3] = {0, 0, 0};
            for (int a = 0; a < 8; a++) {
                const int32_t nd = ien[e * 8 + a] - 1;
                p_ien[a * np + p] = nd;
                if (ndof[nd] == 12) m |= (uint8_t)(1u << a);
                for (int i = 0; i < 3; i++) xc[i] = xc[i] + coor[(int64_t)nd * 3 + i];
            }
            for (int i = 0; i < 3; i++) xc[i] = xc[i] / 8;
            p_mask[p] = m;
            double d[3];
            pml_elem_damping(s, xc, d);
            for (int i = 0; i < 3; i++) p_damp[i * np + p] = d[i];
            for (int i = 0; i < 24; i++) {
                p_shp[i * np + p] = shp[e * 24 + i];
                p_em[i * np + p] = emass[e * 24 + i];
            }
            p_det[p] = det[e];
            p_vp[p] = mat[0 * ne + e];
            p_rho[p] = mat[2 * ne + e];
            p_lam[p] = mat[3 * ne + e];
            p_mu[p] = mat[4 * ne + e];
            for (int i = 0; i < 6; i++) p_ss[i * np + p] = ss[e * 6 + i];
            for (int mm = 0; mm < 4; mm++)
                for (int a = 0; a < 8; a++) p_phi[(mm * 8 + a) * np + p] = phi[e * 32 + mm * 8 + a];
            for (int i = 0; i < 21; i++) p_st[i * np + p] = stress[(int64_t)sci[e] + i];
        }
        int rc;
        int32_t *pi;
        uint8_t *pm;
        if ((rc = dupload(h, &pi, to_blocks<int32_t>(p_ien, np, 8)))) return rc;
        if ((rc = dupload(h, &pm, p_mask))) return rc;
        auto fillp = [&](auto &P) -> int {
            using R = std::remove_pointer_t<decltype(P.ef)>;
            auto up = [&](const R **dst, const std::vector<double> &src, int w) -> int {
                R *q;
                int r2 = dupload(h, &q, to_blocks<R>(src, np, w));
                *dst = q;
                return r2;
            };
            int r2;
            P.n = np;
            P.ien = pi;
            P.mask12 = pm;
            std::vector<uint32_t> cls;
            std::vector<R> tab;
            const int64_t maxc = no_classes ? 0 : std::max<int64_t>(1, np / 16);
            // the damping varies element by element across the layer: it stays a
            // per-element array, so the operator rows (as for regular elements)
            // collapse to a few classes
            if ((r2 = up(&P.damp, p_damp, 3))) return r2;
            if ((r2 = upload_operators<R>(h, P, np, no_classes, [&](int64_t p, R *row) {
                    for (int i = 0; i < 24; i++) row[kOpShp + i] = (R)p_shp[i * np + p];
                    row[kOpDet] = (R)p_det[p];
                    row[kOpLam] = (R)p_lam[p];
                    row[kOpMu] = (R)p_mu[p];
                    row[kOpVp] = (R)p_vp[p];
                    row[kOpRho] = (R)p_rho[p];
                    for (int i = 0; i < 6; i++) row[kOpSs + i] = (R)p_ss[i * np + p];
                }, p_shp, p_phi, &h->op_classes_pml, &h->phi_classes_pml, &h->phi_compact_pml,
                geom_pml, [&](int64_t p, R *row) {
                    const int64_t e = h->pml_ids[p];
                    double xl[3][8], sh[24], s6[6], ph[32];
                    for (int a = 0; a < 8; a++)
                        for (int j = 0; j < 3; j++) xl[j][a] = coor[(int64_t)(ien[e * 8 + a] - 1) * 3 + j];
                    for (int i = 0; i < 24; i++) sh[i] = p_shp[i * np + p];
                    for (int i = 0; i < 6; i++) s6[i] = p_ss[i * np + p];
                    for (int i = 0; i < 32; i++) ph[i] = p_phi[i * np + p];
                    return geometry_row<R>(xl, p_lam[p], p_mu[p], p_vp[p], p_rho[p], s.w, sh, p_det[p], s6, ph, row);
                })))
                return r2;
            if (s.rdampm != 0.0 || s.C_elastic == 0) {
                if ((r2 = up(&P.emass, p_em, 24))) return r2;
            }
            R *q;
            if ((r2 = dupload(h, &q, to_blocks<R>(p_st, np, 21)))) return r2;
            P.stress = q;
            if ((r2 = dalloc(h, &q, blocked_count(np, kEfPml)))) return r2;
            P.ef = q;
            if ((r2 = dalloc(h, &q, blocked_count(np, 6)))) return r2;
            P.ssum = q;
            return 0;
        };
        if (fp32) {
            if ((rc = fillp(h->P32))) return rc;
        } else {
            if ((rc = fillp(h->P64))) return rc;
        }
    }

    // ---------------- node-to-element gather table (ascending element order)
    {
        std::vector<int32_t> cls_index(ne);
        for (int64_t t = 0; t < nr; t++) cls_index[h->reg_ids[t]] = (int32_t)t;
        for (int64_t p = 0; p < np; p++) cls_index[h->pml_ids[p]] = (int32_t)p;
        std::vector<int32_t> rowptr(nn + 1, 0);
        for (int64_t e = 0; e < ne; e++)
            for (int a = 0; a < 8; a++) rowptr[ien[e * 8 + a]]++;  // node id is 1-based: counts land at [node]
        for (int64_t n = 0; n < nn; n++) rowptr[n + 1] += rowptr[n];
        std::vector<int32_t> fillp(rowptr.begin(), rowptr.end() - 1);
        std::vector<uint32_t> ent(rowptr[nn]);
        for (int64_t e = 0; e < ne; e++) {
            const bool pml = etype[e] == 2;
            for (int a = 0; a < 8; a++) {
                const int32_t nd = ien[e * 8 + a] - 1;
                ent[fillp[nd]++] = (pml ? kEntPml : 0u) | ((uint32_t)a << kEntLocalShift) | (uint32_t)cls_index[e];
            }
        }
        for (int64_t q = 0; q < npn; q++) {
            const int32_t n = h->pml_node_ids[q];
            for (int j = rowptr[n]; j < rowptr[n + 1]; j++)
                if (!(ent[j] & kEntPml)) return fail(h, EQD_EINVAL, "PML node %d touches a regular element", n + 1);
        }
        int rc;
        int32_t *d_rp;
        uint32_t *d_ent;
        if ((rc = dupload(h, &d_rp, rowptr))) return rc;
        if ((rc = dupload(h, &d_ent, ent))) return rc;
        h->N.rowptr = d_rp;
        h->N.ent = d_ent;
        // the free 3-component nodes' entries again, slot-major by position,
        // so that a wavefront's load of slot k reads consecutive words
        // (EQD_NODE_ELL=0 keeps the CSR gather)
        const char *ell_env = std::getenv("EQD_NODE_ELL");
        bool ell_ok = !(ell_env && ell_env[0] == '0');
        const int64_t n3 = (int64_t)h->node3_free.size();
        for (int64_t i = 0; i < n3 && ell_ok; i++) {
            const int32_t n = h->node3_free[i];
            if (rowptr[n + 1] - rowptr[n] > kEll3) ell_ok = false;
        }
        h->d_ell3 = nullptr;
        if (ell_ok && n3 > 0) {
            std::vector<uint32_t> ell((size_t)blocked_count(n3, kEll3), kEntNone);
            for (int64_t i = 0; i < n3; i++) {
                const int32_t n = h->node3_free[i];
                for (int j = rowptr[n]; j < rowptr[n + 1]; j++) ell[blocked_index(i, j - rowptr[n], kEll3)] = ent[j];
            }
            if ((rc = dupload(h, &h->d_ell3, ell))) return rc;
        }
    }
    {
        int rc;
        double *d;
        h->N.nn = nn;
        if (s.precision == EQD_PREC_FP32) {
            // FP32 velocity, displacement as the pair hi + lo (eqd_device.h,
            // NodeState); split nodes keep FP64 copies in the fault's arrays
            std::vector<float> hi(3 * nn), m32(nn);
            std::vector<uint16_t> lo(3 * nn);
            for (int64_t j = 0; j < 3 * nn; j++) {
                hi[j] = (float)disp[j];
                lo[j] = bf16_bits((float)(disp[j] - (double)hi[j]));
            }
            for (int64_t n = 0; n < nn; n++) m32[n] = (float)mass[n];
            float *f;
            uint16_t *l;
            if ((rc = dupload(h, &f, to_blocks<float>(vel, nn, 3)))) return rc; h->N.vel32 = f;
            if ((rc = dupload(h, &f, to_blocks<float>(hi, nn, 3)))) return rc; h->N.dhi = f;
            if ((rc = dupload(h, &l, to_blocks<uint16_t>(lo, nn, 3)))) return rc; h->N.dlo = l;
            if ((rc = dupload(h, &f, m32))) return rc; h->N.mass32 = f;
            std::vector<int32_t> slot(nn, -1);
            for (int64_t i = 0; i < s.nf; i++)
                for (int j = 0; j < 2; j++) slot[nsmp[i * 2 + j] - 1] = (int32_t)(2 * i + j);
            int32_t *ps;
            if ((rc = dupload(h, &ps, slot))) return rc; h->N.pair_slot = ps;
        } else {
            if ((rc = dupload(h, &d, to_blocks<double>(vel, nn, 3)))) return rc; h->N.vel = d;
            if ((rc = dupload(h, &d, to_blocks<double>(disp, nn, 3)))) return rc; h->N.disp = d;
        }
        if ((rc = dupload(h, &d, to_blocks<double>(acc, nn, 3)))) return rc; h->N.acc = d;
        if ((rc = dupload(h, &d, mass))) return rc; h->N.mass = d;
        unsigned long long *fl;
        if ((rc = dalloc(h, &fl, 1))) return rc;
        HIPCHK(h, hipMemset(fl, 0xFF, sizeof(unsigned long long)));
        h->N.nanflag = fl;
        if ((rc = dupload(h, &h->d_node3_free, h->node3_free))) return rc;
        if ((rc = dupload(h, &h->d_node3_all, h->node3_all))) return rc;
        PmlNodes &PN = h->PN;
        PN.n = npn;
        int32_t *pn;
        if ((rc = dupload(h, &pn, h->pml_node_ids))) return rc; PN.node = pn;
        if ((rc = dupload(h, &d, to_blocks<double>(pdamp, npn, 3)))) return rc; PN.damp = d;
        if (s.precision == EQD_PREC_FP32) {
            std::vector<float> d32(pdamp.begin(), pdamp.end());
            float *f;
            if ((rc = dupload(h, &f, to_blocks<float>(d32, npn, 3)))) return rc; PN.damp32 = f;
        }
        if (s.precision == EQD_PREC_FP32) {
            float *f;
            if ((rc = dupload(h, &f, to_blocks<float>(pv12, npn, 12)))) return rc; PN.v12f = f;
        } else {
            if ((rc = dupload(h, &d, to_blocks<double>(pv12, npn, 12)))) return rc; PN.v12 = d;
        }
        if ((rc = dupload(h, &d, to_blocks<double>(pacc, npn, 12)))) return rc; PN.acc = d;
    }

    // ---------------- fault
    {
        const int64_t nf = s.nf;
        FaultState &F = h->F;
        F.nf = nf;
        F.ld = (nf + 127) / 128 * 128;  // the fault kernel's grid, lanes past nf write padding rows
        F.nstep = s.nstep;
        const int64_t ld = F.ld;
        std::vector<int32_t> fn(2 * nf), sta(nf, -1);
        std::vector<double> un(3 * nf), us(3 * nf), ud(3 * nf), xyz(3 * nf), fms(nf), fmm(nf), fr(100 * ld, 0.0);
        const double *pun = F64(h, EQD_UN), *pus = F64(h, EQD_US), *pud = F64(h, EQD_UD), *fric = F64(h, EQD_FRIC);
        for (int64_t i = 0; i < nf; i++) {
            fn[i] = nsmp[i * 2] - 1;
            fn[nf + i] = nsmp[i * 2 + 1] - 1;
            for (int k = 0; k < 3; k++) {
                un[k * nf + i] = pun[i * 3 + k];
                us[k * nf + i] = pus[i * 3 + k];
                ud[k * nf + i] = pud[i * 3 + k];
                xyz[k * nf + i] = coor[(int64_t)fn[i] * 3 + k];
            }
            fms[i] = fnms[fn[i]];
            fmm[i] = fnms[fn[nf + i]];
            for (int k = 0; k < 100; k++) fr[k * ld + i] = fric[i * 100 + k];
        }
        if (s.non > 0) {
            const int32_t *an = I32(h, EQD_ANONFS);
            // the reference records station j at every pair with anonfs(1,j) == pair on fault 1
            for (int64_t j = 0; j < s.non; j++) {
                const int32_t pair = an[j * 3], flt = an[j * 3 + 2];
                if (flt != 1) continue;
                if (pair < 1 || pair > nf) return fail(h, EQD_EINVAL, "anonfs(1,%lld) out of range", (long long)j + 1);
                if (sta[pair - 1] >= 0) return fail(h, EQD_EINVAL, "two on-fault stations share node pair %d", pair);
                sta[pair - 1] = (int32_t)j;
            }
        }
        int rc;
        int32_t *pi;
        double *d;
        if (s.precision == EQD_PREC_FP32 && nf > 0) {
            // FP64 velocity and displacement of the split nodes
            std::vector<double> sv(6 * ld, 0.0), su(6 * ld, 0.0);
            for (int64_t i = 0; i < nf; i++)
                for (int side = 0; side < 2; side++)
                    for (int k = 0; k < 3; k++) {
                        const int64_t n = fn[side * nf + i];
                        sv[(side * 3 + k) * ld + i] = vel[k * nn + n];
                        su[(side * 3 + k) * ld + i] = disp[k * nn + n];
                    }
            if ((rc = dupload(h, &d, sv))) return rc; F.sv = d;
            if ((rc = dupload(h, &d, su))) return rc; F.su = d;
        }
        if ((rc = dupload(h, &pi, fn))) return rc; F.nsmp = pi;
        if ((rc = dupload(h, &pi, sta))) return rc; F.station = pi;
        if ((rc = dupload(h, &d, un))) return rc; F.un = d;
        if ((rc = dupload(h, &d, us))) return rc; F.us = d;
        if ((rc = dupload(h, &d, ud))) return rc; F.ud = d;
        if ((rc = dupload(h, &d, std::vector<double>(F64(h, EQD_ARN), F64(h, EQD_ARN) + nf)))) return rc; F.arn = d;
        if ((rc = dupload(h, &d, fms))) return rc; F.fms = d;
        if ((rc = dupload(h, &d, fmm))) return rc; F.fmm = d;
        if ((rc = dupload(h, &d, xyz))) return rc; F.xyz = d;
        if ((rc = dupload(h, &d, fr))) return rc; F.fric = d;
        if ((rc = dupload(h, &pi, std::vector<int32_t>(5 * ld, 0)))) return rc; F.nstat = pi;
        F.k_fallback = 0;
        if ((rc = dupload(h, &d, std::vector<double>(F64(h, EQD_FNFT), F64(h, EQD_FNFT) + nf)))) return rc; F.fnft = d;
        const int64_t non = s.non > 0 ? s.non : 1;
        if ((rc = dalloc(h, &d, non * (int64_t)s.nstep * 12))) return rc;
        HIPCHK(h, hipMemset(d, 0, sizeof(double) * non * s.nstep * 12));
        F.hist = d;
        // hypocentre pairs, as showSourceDynamics finds them
        // (src/faulting.f90:348-349): the slave node within tol = 1e-5 m of
        // xsource along x and of zsource along z
        {
            std::vector<int32_t> hy(nf, -1);
            int32_t nh = 0;
            for (int64_t i = 0; i < nf; i++) {
                const int64_t n = fn[i];
                if (std::fabs(coor[n * 3] - s.xsource) < 1.0e-5 && std::fabs(coor[n * 3 + 2] - s.zsource) < 1.0e-5)
                    hy[i] = nh++;
            }
            F.nhyp = nh;
            F.hyp = nullptr;
            F.hyp_hist = nullptr;
            h->hyp_pairs.clear();
            for (int64_t i = 0; i < nf; i++)
                if (hy[i] >= 0) h->hyp_pairs.push_back((int32_t)i);
            if (nh > 0) {
                int32_t *hp;
                if ((rc = dupload(h, &hp, hy))) return rc;
                F.hyp = hp;
                if ((rc = dalloc(h, &d, (int64_t)s.nstep * nh * kHypVals))) return rc;
                HIPCHK(h, hipMemset(d, 0, sizeof(double) * s.nstep * nh * kHypVals));
                F.hyp_hist = d;
            }
        }
        F.tph = nullptr;
        if (s.friclaw == 5) {
            if ((rc = dalloc(h, &d, 2 * nf * (int64_t)s.nstep))) return rc;
            HIPCHK(h, hipMemset(d, 0, sizeof(double) * 2 * nf * s.nstep));
            F.tph = d;
        }
    }
    // ---------------- off-fault stations
    {
        OffStations &S = h->S;
        S.nrow = 6 * s.noff;
        if (S.nrow > 0) {
            const int32_t *idh = I32(h, EQD_IDHIST);
            std::vector<int32_t> t(3 * S.nrow);
            for (int i = 0; i < S.nrow; i++) {
                t[i] = idh[i * 3] - 1;
                t[S.nrow + i] = idh[i * 3 + 1] - 1;
                t[2 * S.nrow + i] = idh[i * 3 + 2];
                if (t[i] < 0 || t[i] >= nn || t[S.nrow + i] < 0 || t[S.nrow + i] > 2)
                    return fail(h, EQD_EINVAL, "idhist row %d out of range", i + 1);
            }
            int rc;
            int32_t *pi;
            double *d;
            if ((rc = dupload(h, &pi, t))) return rc;
            S.idh = pi;
            const int64_t cnt = (int64_t)(S.nrow + 1) * s.nstep;
            if ((rc = dalloc(h, &d, cnt))) return rc;
            HIPCHK(h, hipMemset(d, 0, sizeof(double) * cnt));
            S.hist = d;
        }
    }
    // ---------------- constants
    StepConsts &C = h->C;
    C.dt = s.dt;
    C.w = s.w;
    C.rdampk = s.rdampk;
    C.rdampm = s.rdampm;
    C.gterm = (1.0 - (double)s.C_elastic) * s.grav * (s.roumax - (s.gamar + 1.0) * s.rhow) / s.roumax;
    C.ccosphi = s.ccosphi;
    C.sinphi = s.sinphi;
    C.exp_dt_tv = s.C_elastic == 0 ? std::exp(-s.dt / s.tv) : 0.0;
    C.kapa_hg = s.kapa_hg;
    C.fric_tp_h = s.fric_tp_h;
    C.pi = s.pi;
    C.slipRateThres = s.slipRateThres;
    C.xsource = s.xsource; C.ysource = s.ysource; C.zsource = s.zsource;
    C.nucR = s.nucR; C.nucRuptVel = s.nucRuptVel; C.nucdtau0 = s.nucdtau0; C.nucT = s.nucT;
    C.C_elastic = s.C_elastic; C.C_Q = s.C_Q; C.C_hg = s.C_hg; C.C_nuclea = s.C_nuclea; C.friclaw = s.friclaw;
    C.TPV = s.TPV; C.insertFaultType = s.insertFaultType; C.nucfault = s.nucfault;
    {
        int rc = dalloc(h, &h->d_clock, 1);
        if (rc) return rc;
        HIPCHK(h, hipMemset(h->d_clock, 0, sizeof(Clock)));
    }
    HIPCHK(h, hipDeviceSynchronize());
    // keep what eqd_get needs to rebuild reference-layout arrays; the caller's
    // buffers are not referenced after this point
    for (int f : {EQD_STRESSARR, EQD_PSTRAIN, EQD_V1, EQD_NODALFORCEARR}) keep_copy(h, f, h->fd);
    for (int f : {EQD_EQNUMINDEXARR, EQD_EQNUMSTARTINDEXLOC, EQD_STRESSCOMPINDEXARR, EQD_NSMP}) keep_copy(h, f, h->fi);
    h->src.clear();
    h->committed = true;
    h->version++;
    return EQD_OK;
}

}  // extern "C"

extern "C" int eqd_run(eqd_handle *h, int32_t nsteps)
{
    if (!h) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_run before eqd_commit");
    if (h->ensemble) return fail(h, EQD_ESTATE, "the handle belongs to an ensemble: run the ensemble");
    if (nsteps < 0) return fail(h, EQD_EINVAL, "nsteps < 0");
    SETTLE(h);
    if (nsteps == 0) return EQD_OK;
    const int rc = engine_run(h->eng, nsteps);
    if (rc) h->err = h->eng->err;
    return rc;
}

extern "C" int eqd_run_async(eqd_handle *h, int32_t nsteps)
{
    if (!h) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_run_async before eqd_commit");
    if (h->ensemble) return fail(h, EQD_ESTATE, "the handle belongs to an ensemble: run the ensemble");
    if (nsteps < 0) return fail(h, EQD_EINVAL, "nsteps < 0");
    SETTLE(h);
    if (nsteps == 0) return EQD_OK;
    const int rc = engine_launch(h->eng, nsteps);
    if (rc) {
        h->err = h->eng->err;
        // whatever was enqueued finishes before the error is returned
        (void)hipStreamSynchronize(h->eng->st);
        (void)hipStreamSynchronize(h->eng->st2);
        return rc;
    }
    h->inflight = nsteps;
    return EQD_OK;
}

extern "C" int eqd_wait(eqd_handle *h)
{
    if (!h) return EQD_EINVAL;
    return settle(h);
}

template <typename T>
static int d2h(eqd_handle *h, std::vector<T> &v, const T *d, int64_t count)
{
    v.resize(count);
    if (count > 0) HIPCHK(h, hipMemcpy(v.data(), d, sizeof(T) * count, hipMemcpyDeviceToHost));
    return 0;
}

// Velocities (vel) or displacements of every node in blocked layout
// (blk_at<3>) at the precision the device keeps: the FP64 arrays, or in the
// FP32 bulk mode the FP32 values (displacement hi + lo) with the split nodes'
// FP64 state from the fault's arrays.
static int node_state(eqd_handle *h, bool vel, std::vector<double> &out)
{
    const int64_t nn = h->s.nn;
    int rc;
    if (!h->N.vel32) return d2h(h, out, vel ? h->N.vel : h->N.disp, blocked_count(nn, 3));
    std::vector<float> a;
    std::vector<uint16_t> b;
    if ((rc = d2h(h, a, vel ? h->N.vel32 : h->N.dhi, blocked_count(nn, 3)))) return rc;
    out.assign(a.begin(), a.end());
    if (!vel) {
        if ((rc = d2h(h, b, h->N.dlo, blocked_count(nn, 3)))) return rc;
        for (size_t j = 0; j < out.size(); j++) out[j] = (double)a[j] + (double)bf16_value(b[j]);
    }
    const int64_t nf = h->F.nf, ld = h->F.ld;
    if (nf > 0) {
        std::vector<double> st;
        if ((rc = d2h(h, st, vel ? h->F.sv : h->F.su, 6 * ld))) return rc;
        const int32_t *nsmp = I32(h, EQD_NSMP);
        for (int64_t i = 0; i < nf; i++)
            for (int side = 0; side < 2; side++)
                for (int k = 0; k < 3; k++)
                    out[blocked_index(nsmp[i * 2 + side] - 1, k, 3)] = st[(side * 3 + k) * ld + i];
    }
    return EQD_OK;
}

extern "C" {

int eqd_get(eqd_handle *h, int32_t field, void *host, int64_t count)
{
    if (!h) return EQD_EINVAL;
    SETTLE(h);
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_get before eqd_commit");
    FieldSpec sp;
    if (!field_spec(h->s, field, sp)) return fail(h, EQD_EINVAL, "unknown field %d", field);
    if (count != sp.count) return fail(h, EQD_EINVAL, "field %d: count %lld, expected %lld", field, (long long)count, (long long)sp.count);
    HIPCHK(h, hipSetDevice(h->s.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const eqd_scalars &s = h->s;
    const int64_t nn = s.nn, nf = s.nf;
    double *out = (double *)host;
    int rc;
    switch (field) {
    case EQD_VELARR:
    case EQD_DISPARR: {
        std::vector<double> v;
        if ((rc = node_state(h, field == EQD_VELARR, v))) return rc;
        for (int64_t n = 0; n < nn; n++)
            for (int k = 0; k < 3; k++) out[n * 3 + k] = v[blocked_index(n, k, 3)];
        return EQD_OK;
    }
    case EQD_V1:
    case EQD_NODALFORCEARR: {
        // v1 of 3-component nodes is velArr; of PML nodes the 12 split velocities
        const bool isv = field == EQD_V1;
        std::vector<double> v, p;
        if (isv) {
            if ((rc = node_state(h, true, v))) return rc;
        } else if ((rc = d2h(h, v, h->N.acc, blocked_count(nn, 3)))) {
            return rc;
        }
        if (isv && h->PN.v12f) {
            std::vector<float> f;
            if ((rc = d2h(h, f, h->PN.v12f, blocked_count(h->PN.n, 12)))) return rc;
            p.assign(f.begin(), f.end());
        } else if ((rc = d2h(h, p, isv ? h->PN.v12 : h->PN.acc, blocked_count(h->PN.n, 12)))) {
            return rc;
        }
        const auto &orig = h->fd[field];  // values for equations the device does not own (fixed PML nodes)
        for (int64_t i = 0; i < s.neq; i++) out[i] = orig.empty() ? 0.0 : orig[i];
        const int32_t *eqi = I32(h, EQD_EQNUMINDEXARR), *esl = I32(h, EQD_EQNUMSTARTINDEXLOC);
        for (int32_t n : h->node3_all)
            for (int k = 0; k < 3; k++) out[eqi[esl[n] + k] - 1] = v[blocked_index(n, k, 3)];
        const int64_t np = h->PN.n;
        for (int64_t q = 0; q < np; q++) {
            const int32_t n = h->pml_node_ids[q];
            for (int k = 0; k < 12; k++) out[eqi[esl[n] + k] - 1] = p[blocked_index(q, k, 12)];
        }
        return EQD_OK;
    }
    case EQD_FRIC: {
        std::vector<double> v;
        const int64_t ld = h->F.ld;
        if ((rc = d2h(h, v, h->F.fric, 100 * ld))) return rc;
        for (int64_t i = 0; i < nf; i++)
            for (int k = 0; k < 100; k++) out[i * 100 + k] = v[k * ld + i];
        return EQD_OK;
    }
    case EQD_FNFT: {
        if (nf > 0) HIPCHK(h, hipMemcpy(out, h->F.fnft, sizeof(double) * nf, hipMemcpyDeviceToHost));
        return EQD_OK;
    }
    case EQD_ONFAULTHIST: {
        if (count > 0) HIPCHK(h, hipMemcpy(out, h->F.hist, sizeof(double) * count, hipMemcpyDeviceToHost));
        return EQD_OK;
    }
    case EQD_OFFFAULTHIST: {
        if (count > 0) HIPCHK(h, hipMemcpy(out, h->S.hist, sizeof(double) * count, hipMemcpyDeviceToHost));
        return EQD_OK;
    }
    case EQD_STRESSARR: {
        const auto &orig = h->fd[EQD_STRESSARR];
        for (int64_t i = 0; i < s.size_stress; i++) out[i] = orig[i];
        const int32_t *sci = I32(h, EQD_STRESSCOMPINDEXARR);
        const int64_t nr = h->n_reg, np = h->n_pml;
        std::vector<double> v;
        if (s.precision == EQD_PREC_FP32) {
            std::vector<float> f;
            if ((rc = d2h(h, f, h->E32.stress, blocked_count(nr, h->nreg_stress)))) return rc;
            v.assign(f.begin(), f.end());
        } else if ((rc = d2h(h, v, h->E64.stress, blocked_count(nr, h->nreg_stress)))) {
            return rc;
        }
        for (int64_t t = 0; t < nr; t++) {
            const int64_t e = h->reg_ids[t];
            for (int i = 0; i < h->nreg_stress; i++) {
                const int64_t idx = (int64_t)sci[e] + i;
                if (idx < s.size_stress) out[idx] = v[blocked_index(t, i, h->nreg_stress)];
            }
        }
        if (s.precision == EQD_PREC_FP32) {
            std::vector<float> f;
            if ((rc = d2h(h, f, h->P32.stress, blocked_count(np, 21)))) return rc;
            v.assign(f.begin(), f.end());
        } else if ((rc = d2h(h, v, h->P64.stress, blocked_count(np, 21)))) {
            return rc;
        }
        for (int64_t p = 0; p < np; p++) {
            const int64_t e = h->pml_ids[p];
            for (int i = 0; i < 21; i++) out[(int64_t)sci[e] + i] = v[blocked_index(p, i, 21)];
        }
        return EQD_OK;
    }
    case EQD_PSTRAIN: {
        const auto &orig = h->fd[EQD_PSTRAIN];
        for (int64_t i = 0; i < s.ne; i++) out[i] = orig[i];
        std::vector<double> v;
        const double *src = s.precision == EQD_PREC_FP32 ? h->E32.pstrain : h->E64.pstrain;
        if ((rc = d2h(h, v, src, h->n_reg))) return rc;
        for (int64_t t = 0; t < h->n_reg; t++) out[h->reg_ids[t]] = v[t];
        return EQD_OK;
    }
    default:
        return fail(h, EQD_EINVAL, "field %d is an input and is not kept after eqd_commit", field);
    }
}

int eqd_clock(eqd_handle *h, int32_t *nt, double *time_elapsed)
{
    if (!h) return EQD_EINVAL;
    SETTLE(h);
    if (nt) *nt = h->nt;
    if (time_elapsed) *time_elapsed = h->time;
    return EQD_OK;
}

int eqd_nan_info(eqd_handle *h, int32_t *node, int32_t *nt)
{
    if (!h || !h->committed) return EQD_EINVAL;
    (void)settle(h);  // the NaN a run in flight found is reported below
    unsigned long long key;
    HIPCHK(h, hipMemcpy(&key, h->N.nanflag, sizeof(key), hipMemcpyDeviceToHost));
    const bool none = key == ~0ull;
    if (node) *node = none ? 0 : (int32_t)(uint32_t)(key & 0xFFFFFFFFu);
    if (nt) *nt = none ? 0 : (int32_t)(uint32_t)(key >> 32);
    return EQD_OK;
}

int eqd_kernel_time(eqd_handle *h, const char *name, double *avg_ms, int64_t *launches)
{
    if (!h || !name) return EQD_EINVAL;
    SETTLE(h);
    if (!h->eng) return fail(h, EQD_ESTATE, "eqd_kernel_time before eqd_commit");
    auto it = h->eng->timers.find(name);
    if (it == h->eng->timers.end() || it->second.launches == 0) {
        if (avg_ms) *avg_ms = 0.0;
        if (launches) *launches = 0;
        return EQD_OK;
    }
    if (avg_ms) *avg_ms = it->second.total_ms / (double)it->second.launches;
    if (launches) *launches = it->second.launches;
    return EQD_OK;
}

int eqd_set_timing(eqd_handle *h, int32_t on)
{
    if (!h) return EQD_EINVAL;
    SETTLE(h);
    h->timing = on < 0 ? 0 : on;
    if (h->eng) h->eng->timing = h->timing;
    return EQD_OK;
}

int64_t eqd_device_bytes(eqd_handle *h) { return h ? h->bytes : 0; }

int eqd_set_guess(eqd_handle *h, const float *params, int64_t nparams, int32_t k_fallback, float gate)
{
    if (!h) return EQD_EINVAL;
    SETTLE(h);
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_set_guess before eqd_commit");
    if (k_fallback < 0 || k_fallback > 20) return fail(h, EQD_EINVAL, "k_fallback %d outside 0..20", k_fallback);
    if (!params || nparams == 0) {
        // plain guess: the reference's solve (k_fallback 0) or the faster
        // mode's iteration policy without the network (k_fallback > 0)
        h->learn = false;
        h->F.k_fallback = k_fallback;
        h->version++;
        return EQD_OK;
    }
    if (nparams != EQD_MLP_NPARAMS)
        return fail(h, EQD_EINVAL, "the network has %d parameters, got %lld", EQD_MLP_NPARAMS, (long long)nparams);
    if (k_fallback < 1 || k_fallback > 20) return fail(h, EQD_EINVAL, "k_fallback %d outside 1..20", k_fallback);
    if (!(gate >= 0.0f) || !std::isfinite(gate)) return fail(h, EQD_EINVAL, "gate must be finite and >= 0");
    if (h->s.friclaw < 3) return fail(h, EQD_EINVAL, "the learned guess needs rate-and-state friction (friclaw 3, 4)");
    for (int64_t k = 0; k < nparams; k++)
        if (!std::isfinite(params[k])) return fail(h, EQD_EINVAL, "parameter %lld is not finite", (long long)k);
    if (params[kOffNorm + 2 * kMlpIn + 1] == 0.0f) return fail(h, EQD_EINVAL, "out_scale is zero");
    HIPCHK(h, hipSetDevice(h->s.device));
    int rc;
    if (!h->mlp_p) {
        if ((rc = dalloc(h, &h->mlp_p, (int64_t)kMlpNParams))) return rc;
        if ((rc = dalloc(h, &h->mlp_wb, (int64_t)(64 * 16 + 4 * 4096)))) return rc;
        if ((rc = dalloc(h, &h->mlp_w4r, (int64_t)64))) return rc;
    }
    HIPCHK(h, hipMemcpyAsync(h->mlp_p, params, sizeof(float) * kMlpNParams, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, launch_mlp_prepare(h->mlp_p, h->mlp_wb, h->mlp_w4r, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    MlpNet &M = h->F.net;
    M.w1 = h->mlp_wb;
    M.w2 = h->mlp_wb + 64 * 16;
    M.w3 = M.w2 + 4096;
    M.w2t = M.w3 + 4096;
    M.w3t = M.w2t + 4096;
    M.b1 = h->mlp_p + kOffB1;
    M.b2 = h->mlp_p + kOffB2;
    M.b3 = h->mlp_p + kOffB3;
    M.w4r = h->mlp_w4r;
    M.b4 = h->mlp_p + kOffB4;
    M.norm = h->mlp_p + kOffNorm;
    h->F.k_fallback = k_fallback;
    h->F.gate = gate;
    h->learn = true;
    h->version++;
    return EQD_OK;
}

int eqd_harvest(eqd_handle *h, int32_t every, int64_t capacity)
{
    if (!h) return EQD_EINVAL;
    SETTLE(h);
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_harvest before eqd_commit");
    if (capacity < 0 || (capacity > 0 && every < 1)) return fail(h, EQD_EINVAL, "bad harvest every/capacity");
    HIPCHK(h, hipSetDevice(h->s.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (h->harvest) {
        (void)hipFree(h->harvest);
        h->harvest = nullptr;
    }
    h->F.harvest = nullptr;
    h->F.harvest_cap = 0;
    h->F.harvest_every = 0;
    h->F.harvest_t0 = h->nt;
    h->version++;
    if (capacity == 0) return EQD_OK;
    if (hipMalloc(&h->harvest, sizeof(float) * 7 * (size_t)capacity) != hipSuccess) {
        h->harvest = nullptr;
        return fail(h, EQD_EOOM, "harvest buffer of %lld rows", (long long)capacity);
    }
    h->F.harvest = h->harvest;
    h->F.harvest_cap = capacity;
    h->F.harvest_every = every;
    return EQD_OK;
}

int eqd_harvest_get(eqd_handle *h, float *out, int64_t max_rows, int64_t *rows)
{
    if (!h || !rows || max_rows < 0 || (max_rows > 0 && !out)) return EQD_EINVAL;
    SETTLE(h);
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_harvest_get before eqd_commit");
    HIPCHK(h, hipSetDevice(h->s.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // the fault kernel writes nf rows at every harvest step since harvest_t0
    const FaultState &F = h->F;
    int64_t have = 0;
    if (F.harvest && F.harvest_every > 0)
        have = std::min(F.harvest_cap, F.nf * (int64_t)(h->nt / F.harvest_every - F.harvest_t0 / F.harvest_every));
    *rows = have;
    const int64_t n = std::min(max_rows, have);
    if (n > 0) HIPCHK(h, hipMemcpy(out, h->harvest, sizeof(float) * 7 * n, hipMemcpyDeviceToHost));
    return EQD_OK;
}

static void free_records(eqd_handle *h)
{
    RecStream &q = h->rq;
    if (q.cst) (void)hipStreamSynchronize(q.cst);
    for (const RecSeg &g : q.segs)
        if (g.ev) (void)hipEventDestroy(g.ev);
    for (hipEvent_t ev : q.spare) (void)hipEventDestroy(ev);
    if (q.cst) (void)hipStreamDestroy(q.cst);
    if (q.hvel) (void)hipHostFree(q.hvel);
    if (q.hslip) (void)hipHostFree(q.hslip);
    q = RecStream{};
    for (void *&p : h->rec_bufs) {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    h->d_rec = nullptr;
    h->rec = Recorder{};
}

int eqd_record(eqd_handle *h, const int32_t *node_ids, int32_t n_nodes, int32_t every, int32_t capacity)
{
    if (!h) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_record before eqd_commit");
    SETTLE(h);
    if (n_nodes < 0 || (n_nodes > 0 && !node_ids) || capacity < 0 || (capacity > 0 && every < 1))
        return fail(h, EQD_EINVAL, "bad record nodes/every/capacity");
    for (int32_t i = 0; i < n_nodes; i++)
        if (node_ids[i] < 1 || node_ids[i] > h->s.nn) return fail(h, EQD_EINVAL, "record node %d out of range", i + 1);
    HIPCHK(h, hipSetDevice(h->s.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    free_records(h);
    h->version++;
    if (capacity == 0) return EQD_OK;
    std::vector<int32_t> ids(node_ids, node_ids + n_nodes);
    for (int32_t &v : ids) v -= 1;
    const size_t nv = sizeof(double) * 3 * (size_t)capacity * (size_t)std::max(n_nodes, 1);
    const size_t ns = sizeof(double) * (size_t)capacity * (size_t)std::max<int64_t>(h->F.nf, 1);
    RecStream &q = h->rq;
    if (hipMalloc(&h->rec_bufs[0], sizeof(Recorder) + sizeof(int32_t) * std::max(n_nodes, 1)) != hipSuccess ||
        hipMalloc(&h->rec_bufs[1], nv) != hipSuccess || hipMalloc(&h->rec_bufs[2], ns) != hipSuccess ||
        hipHostMalloc((void **)&q.hvel, nv, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&q.hslip, ns, hipHostMallocDefault) != hipSuccess) {
        free_records(h);
        return fail(h, EQD_EOOM, "record buffers of %d records", capacity);
    }
    // the copies run on a stream of their own (default priority), beside the steps
    HIPCHK(h, hipStreamCreateWithFlags(&q.cst, hipStreamNonBlocking));
    Recorder &R = h->rec;
    R.every = every;
    R.capacity = capacity;
    R.n_nodes = n_nodes;
    R.nodes = (const int32_t *)((char *)h->rec_bufs[0] + sizeof(Recorder));
    R.vel = (double *)h->rec_bufs[1];
    R.slip_rate = h->F.nf > 0 ? (double *)h->rec_bufs[2] : nullptr;
    q.drained = q.next = next_record(h->nt, every);  // the first record is the next record step
    if (n_nodes > 0)
        HIPCHK(h, hipMemcpy((void *)R.nodes, ids.data(), sizeof(int32_t) * n_nodes, hipMemcpyHostToDevice));
    h->d_rec = (Recorder *)h->rec_bufs[0];
    HIPCHK(h, hipMemcpy(h->d_rec, &R, sizeof(Recorder), hipMemcpyHostToDevice));
    return EQD_OK;
}

int eqd_record_get(eqd_handle *h, double *vel, double *slip_rate, int32_t max_records, int32_t *count)
{
    if (!h || !count || max_records < 0) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_record_get before eqd_commit");
    *count = 0;
    if (!h->d_rec) return EQD_OK;
    HIPCHK(h, hipSetDevice(h->s.device));
    RecStream &q = h->rq;
    const Recorder &R = h->rec;
    // every finished run's records; a run still in flight keeps its own
    const size_t nseg = q.segs.size() - (h->inflight > 0 && !q.segs.empty() ? 1 : 0);
    int32_t n = 0, dropped = 0;
    for (size_t i = 0; i < nseg; i++) {
        n += q.segs[i].end - q.segs[i].begin;
        dropped += q.segs[i].next - q.segs[i].end;
    }
    if (n > max_records) return fail(h, EQD_EINVAL, "%d records waiting, room for %d", n, max_records);
    const int64_t nv = 3 * (int64_t)R.n_nodes, nf = h->F.nf;
    int32_t k = 0;
    for (size_t i = 0; i < nseg; i++) {
        RecSeg g = q.segs.front();
        q.segs.pop_front();
        if (g.ev) {
            const hipError_t ec = hipEventSynchronize(g.ev);
            q.spare.push_back(g.ev);
            if (ec != hipSuccess) return fail(h, EQD_EHIP, "record copy: %s", hipGetErrorString(ec));
        }
        for (int32_t r = g.begin; r < g.end; r++, k++) {
            const int64_t a = r % R.capacity;
            if (vel && nv > 0) memcpy(vel + k * nv, q.hvel + a * nv, sizeof(double) * nv);
            if (slip_rate && nf > 0) memcpy(slip_rate + k * nf, q.hslip + a * nf, sizeof(double) * nf);
        }
        q.drained = g.next;
    }
    *count = n;
    if (dropped > 0)
        return fail(h, EQD_EINVAL, "%d record steps since the last eqd_record_get, capacity %d: the last %d dropped",
                    n + dropped, R.capacity, dropped);
    return EQD_OK;
}

int eqd_hypocentre_pairs(eqd_handle *h, int32_t *pairs, int32_t cap, int32_t *n)
{
    if (!h || !n || cap < 0 || (cap > 0 && !pairs)) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_hypocentre_pairs before eqd_commit");
    *n = (int32_t)h->hyp_pairs.size();
    for (int32_t k = 0; k < *n && k < cap; k++) pairs[k] = h->hyp_pairs[k] + 1;
    return EQD_OK;
}

int eqd_hypocentre_get(eqd_handle *h, int32_t first_step, int32_t count, double *out)
{
    if (!h || count < 0 || (count > 0 && !out)) return EQD_EINVAL;
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_hypocentre_get before eqd_commit");
    if (count == 0 || h->F.nhyp == 0) return EQD_OK;
    // only steps of runs that have ended: their rows are final, whatever a
    // run in flight writes to the rows of later steps
    const int64_t last = (int64_t)first_step + count - 1;
    if (first_step < 1 || last > h->nt)
        return fail(h, EQD_EINVAL, "steps %d..%lld are not all done (clock %d)", first_step, (long long)last, h->nt);
    HIPCHK(h, hipSetDevice(h->s.device));
    const size_t row = (size_t)h->F.nhyp * kHypVals;
    HIPCHK(h, hipMemcpy(out, h->F.hyp_hist + (size_t)(first_step - 1) * row, sizeof(double) * row * count,
                        hipMemcpyDeviceToHost));
    return EQD_OK;
}

int eqd_newton_stats(eqd_handle *h, int64_t stats[5])
{
    if (!h || !stats) return EQD_EINVAL;
    SETTLE(h);
    if (!h->committed) return fail(h, EQD_ESTATE, "eqd_newton_stats before eqd_commit");
    HIPCHK(h, hipSetDevice(h->s.device));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < 5; k++) stats[k] = 0;
    const int64_t nf = h->F.nf, ld = h->F.ld;
    if (nf == 0) return EQD_OK;
    std::vector<int32_t> v;
    int rc;
    if ((rc = d2h(h, v, (const int32_t *)h->F.nstat, 5 * ld))) return rc;
    for (int k = 0; k < 5; k++)
        for (int64_t i = 0; i < nf; i++) stats[k] += v[k * ld + i];
    return EQD_OK;
}

int eqd_layout_info(eqd_handle *h, const char *key, int64_t *value)
{
    if (!h || !key || !value) return EQD_EINVAL;
    const std::string k = key;
    if (k == "regular_elements") *value = h->n_reg;
    else if (k == "pml_elements") *value = h->n_pml;
    else if (k == "op_classes_regular") *value = h->op_classes_reg;
    else if (k == "phi_classes_regular") *value = h->phi_classes_reg;
    else if (k == "op_classes_pml") *value = h->op_classes_pml;
    else if (k == "phi_classes_pml") *value = h->phi_classes_pml;
    else if (k == "phi_compact_regular") *value = h->phi_compact_reg;
    else if (k == "phi_compact_pml") *value = h->phi_compact_pml;
    else if (k == "geom_ops_regular") *value = h->s.precision == EQD_PREC_FP32 ? h->E32.opb_geom : 0;
    else if (k == "geom_ops_pml") *value = h->s.precision == EQD_PREC_FP32 ? h->P32.opb_geom : 0;
    else if (k == "node3_slot_major") *value = h->d_ell3 ? 1 : 0;
    else return fail(h, EQD_EINVAL, "unknown layout key %s", key);
    return EQD_OK;
}

const char *eqd_last_error(eqd_handle *h) { return h ? h->err.c_str() : "null handle"; }

int eqd_destroy(eqd_handle *h)
{
    if (!h) return EQD_OK;
    if (h->ensemble) return fail(h, EQD_ESTATE, "the handle belongs to an ensemble: destroy the ensemble first");
    (void)settle(h);
    (void)hipSetDevice(h->s.device);
    if (h->eng) {
        engine_free(h->eng);
        delete h->eng;
    }
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->harvest) (void)hipFree(h->harvest);
    free_records(h);
    delete h;
    return EQD_OK;
}

}  // extern "C"
