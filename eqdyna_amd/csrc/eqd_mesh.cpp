// Structured case builder: the reference host's mesh generation and one-time
// element precompute, for one rank and one fault: planar vertical at y = 0,
// planar dipping on wedge elements (C_degen > 3) or rough (insertFaultType > 0).
//
// It produces the same globalvar arrays, in the same layout and bit for bit,
// as src/meshgen.f90 (nodes, equation numbers, elements, split nodes, fault
// areas, stations) followed by src/assembleGlobalMass.f90 (shape-function
// derivatives, Jacobians, lumped masses, hourglass operators), so that cases
// far larger than the Fortran host can mesh in reasonable time can be handed
// to the GPU time step.  Element-local precompute runs in parallel (OpenMP);
// everything that accumulates across elements runs in the reference's order.
//
// Two calls: eqdm_size() counts, the caller allocates, eqdm_fill() fills.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "eqd_precompute.h"

extern "C" {

typedef struct {
    // bModelGeometry.txt / bFaultGeometry.txt / bGlobal.txt / bMaterial.txt
    double xmin, xmax, ymin, ymax, zmin, zmax;
    int32_t dis4uniF, dis4uniB;
    double rat, dx, dy, dz;
    double fxmin, fxmax, fymin, fymax, fzmin, fzmax;
    double fstrike;      // degrees
    int32_t nPML;
    int32_t C_elastic;
    double roumax, rhow, gamar, grav;
    double str1ToFaultAngle;  // degrees
    double devStrToStrVertRatio;
    int32_t nmat, n2mat;
    const double *material;  // (nmat, n2mat) row-major (material(i,:) in row i)
    int32_t n_on;            // on-fault stations (x, z) in metres
    const double *st_on;
    int32_t n_off;           // off-fault stations (x, y, z) in metres
    const double *st_off;
    double tol;
    int32_t nthreads;
    // dipping fault on wedges: dip in degrees when > 3 (src/readInputFiles.f90:140-145)
    double C_degen;
    // rough fault: bFault_Rough_Geometry.txt (src/readInputFiles.f90:268-312);
    // rough_geo(3, nx*nz) = (y offset, dy/dx, dy/dz), x-major, z fastest
    int32_t insertFaultType;
    int32_t rough_nx, rough_nz;
    double rough_dx, rough_fxmin, rough_fzmin;
    const double *rough_geo;
} eqdm_params;

typedef struct {
    int32_t nx, ny, nz;
    int64_t nn, ne, neq, size_eqidx, size_stress, nf, non, noff;
    double PMLb[8];
    double xmin, xmax, ymin, ymax, zmin, zmax;
} eqdm_sizes;

typedef struct {
    double *meshCoor;          // (3, nn)
    int32_t *eqNumIndexArr;    // (size_eqidx)
    int32_t *eqNumStartIndexLoc, *numOfDofPerNodeArr;  // (nn)
    int32_t *nodeElemIdRelation;  // (8, ne)
    int32_t *elemTypeArr;      // (ne)
    double *mat;               // (ne, 5)
    double *eleshp;            // (3, 8, ne)
    double *eledet;            // (ne)
    double *elemass;           // (24, ne)
    double *ss;                // (6, ne)
    double *phi;               // (8, 4, ne)
    int32_t *stressCompIndexArr;  // (ne)
    double *stressArr;         // (size_stress)
    double *nodalMassArr;      // (neq)
    double *fnms;              // (nn)
    int32_t *nsmp;             // (2, nf)
    double *un, *us, *ud;      // (3, nf)
    double *arn;               // (nf)
    int32_t *anonfs;           // (3, max(non,1))
    int32_t *idhist;           // (3, 6*noff)
} eqdm_arrays;

}  // extern "C"

namespace {

// getLocalOneDimCoorArrAndSize for one rank, src/meshgen.f90:447-546
static std::vector<double> axis(const eqdm_params &p, int dim, double PMLb[8], double bounds[2])
{
    int nuni;
    double h, front, back, mn, mx;
    if (dim == 1) {
        nuni = (int)std::lround((p.fxmax - p.fxmin) / p.dx) + 1;
        h = p.dx; front = p.fxmin; back = p.fxmax; mn = p.xmin; mx = p.xmax;
    } else if (dim == 2) {
        nuni = p.dis4uniF + p.dis4uniB + 1;
        h = p.dy; front = -(p.dis4uniF * p.dy); back = p.dis4uniB * p.dy; mn = p.ymin; mx = p.ymax;
    } else {
        nuni = (int)std::lround((p.fzmax - p.fzmin) / p.dz) + 1;
        h = p.dz; front = p.fzmin; back = p.fzmax; mn = p.zmin; mx = p.zmax;
    }
    const int np = 1000000;
    double c = front, g = h;
    int i;
    for (i = 1; i <= np; i++) {
        g = g * p.rat;
        c = c - g;
        if (c <= mn) break;
    }
    const int frontId = i + p.nPML;
    c = back;
    g = h;
    for (i = 1; i <= np; i++) {
        g = g * p.rat;
        c = c + g;
        if (c >= mx) break;
    }
    if (dim == 3) i = -p.nPML;
    const int size = nuni + frontId + i + p.nPML;
    std::vector<double> a(size + 1);  // 1-based
    a[frontId + 1] = front;
    g = h;
    for (int k = frontId; k >= 1; k--) {
        g = g * p.rat;
        a[k] = a[k + 1] - g;
    }
    for (int k = frontId + 2; k <= frontId + nuni; k++) a[k] = a[k - 1] + h;
    if (dim < 3) {
        g = h;
        for (int k = frontId + nuni + 1; k <= size; k++) {
            g = g * p.rat;
            a[k] = a[k - 1] + g;
        }
    }
    bounds[0] = a[1];
    bounds[1] = a[size];
    if (dim == 1) {
        PMLb[0] = a[size - p.nPML];
        PMLb[1] = a[p.nPML + 1];
        PMLb[5] = a[size] - a[size - 1];
    } else if (dim == 2) {
        PMLb[2] = a[size - p.nPML];
        PMLb[3] = a[p.nPML + 1];
        PMLb[6] = a[size] - a[size - 1];
    } else {
        PMLb[4] = a[p.nPML + 1];
        PMLb[7] = a[2] - a[1];
    }
    return a;
}

struct Grid {
    std::vector<double> x, y, z;  // 1-based
    int nx, ny, nz;
    double PMLb[8];
    double xmin, xmax, ymin, ymax, zmin, zmax;
};

static Grid make_grid(const eqdm_params &p)
{
    Grid G;
    double b[2];
    G.x = axis(p, 1, G.PMLb, b);
    G.xmin = b[0]; G.xmax = b[1];
    G.y = axis(p, 2, G.PMLb, b);
    G.ymin = b[0]; G.ymax = b[1];
    G.z = axis(p, 3, G.PMLb, b);
    G.zmin = b[0]; G.zmax = b[1];
    G.nx = (int)G.x.size() - 1;
    G.ny = (int)G.y.size() - 1;
    G.nz = (int)G.z.size() - 1;
    return G;
}

static const double kPi = 4.0 * std::atan(1.0);

// checkIsOnFault, src/meshgen.f90:771-786: the fault box, then y == 0 for a
// vertical fault or distance to the plane z = -y*tan(dip) below dx/100
static inline bool on_fault(const eqdm_params &p, double x, double y, double z)
{
    if (!(x >= p.fxmin - p.tol && x <= p.fxmax + p.tol && y >= p.fymin - p.tol && y <= p.fymax + p.tol &&
          z >= p.fzmin - p.tol && z <= p.fzmax + p.tol))
        return false;
    if (p.C_degen == 0.0) return y == 0.0;
    if (p.C_degen > 3.0) {
        const double tg = std::tan(p.C_degen / 180.0 * kPi);
        const double d = std::fabs(z + y * tg) / std::sqrt(1.0 + tg * tg);
        return d < p.dx / 100.0;
    }
    return false;
}

// insertFaultInterface, src/meshgen.f90:921-962: the y coordinate of a node
// column shifted by the rough surface, and the surface slopes there
static inline void rough_shift(const eqdm_params &p, double ymin, double ymax, double x, double y, double z,
                               double *yc, double *pfx, double *pfz)
{
    const double fx1 = p.rough_fxmin, fz1 = p.rough_fzmin;
    const double fx2 = (p.rough_nx - 1) * p.rough_dx + fx1;
    int ixx = 1, izz = 1;
    const bool zin = z > fz1 - p.tol;
    if (x < fx2 + p.tol && x > fx1 - p.tol)
        ixx = (int)std::lround((x - fx1) / p.dx) + 1;
    else if (x > fx2 + p.tol)
        ixx = p.rough_nx;
    if (zin) izz = (int)std::lround((z - fz1) / p.dz) + 1;
    const double *r = p.rough_geo + 3 * ((int64_t)p.rough_nz * (ixx - 1) + izz - 1);
    const double peak = r[0];
    *pfx = r[1];
    *pfz = r[2];
    if (y > -p.tol)
        *yc = y * (ymax - peak) / ymax + peak;
    else
        *yc = y * (peak - ymin) / (-ymin) + peak;
}

static inline bool is_pml_node(const Grid &G, double x, double y, double z)
{
    return x > G.PMLb[0] || x < G.PMLb[1] || y > G.PMLb[2] || y < G.PMLb[3] || z < G.PMLb[4];
}

static inline bool is_fixed(const eqdm_params &p, const Grid &G, double x, double y, double z)
{
    return std::fabs(x - G.xmin) < p.tol || std::fabs(x - G.xmax) < p.tol || std::fabs(y - G.ymin) < p.tol ||
           std::fabs(y - G.ymax) < p.tol || std::fabs(z - G.zmin) < p.tol;
}

struct Counts {
    int64_t nn = 0, ne = 0, neq = 0, size_eqidx = 0, size_stress = 0, nf = 0, non = 0, noff = 0;
};

// off-fault station match of setSurfaceStation, src/meshgen.f90:574-658
static int match_station(const eqdm_params &p, const Grid &G, int ix, int iy, double x, double y, double z,
                         std::vector<char> &found)
{
    const int nx = G.nx, ny = G.ny;
    if (!(iy > 1 && iy < ny)) return -1;
    const int part = (ix > 1 && ix < nx) ? 1 : (ix == 1 ? 2 : (ix == nx ? 3 : 0));
    if (!part) return -1;
    for (int i = 0; i < p.n_off; i++) {
        if (found[i]) continue;
        const double xs = p.st_off[i * 3], ys = p.st_off[i * 3 + 1], zs = p.st_off[i * 3 + 2];
        if (!(std::fabs(z - zs) < p.tol)) continue;
        bool xok;
        if (part == 1)
            xok = std::fabs(x - xs) < p.tol ||
                  (xs > G.x[ix - 1] && xs < x && (x - xs) < (xs - G.x[ix - 1])) ||
                  (xs > x && xs < G.x[ix + 1] && (xs - x) < (G.x[ix + 1] - xs));
        else if (part == 2)
            xok = std::fabs(x - xs) < p.tol || (xs > x && xs < G.x[ix + 1] && (xs - x) < (G.x[ix + 1] - xs));
        else
            xok = xs > G.x[ix - 1] && xs < x && (x - xs) < (xs - G.x[ix - 1]);
        if (!xok) continue;
        const bool yok = std::fabs(y - ys) < p.tol ||
                         (ys > G.y[iy - 1] && ys < y && (y - ys) < (ys - G.y[iy - 1])) ||
                         (ys > y && ys < G.y[iy + 1] && (ys - y) < (G.y[iy + 1] - ys));
        if (!yok) continue;
        found[i] = 1;
        return i;
    }
    return -1;
}

// mat(e, 1:5) = vp, vs, rho, lambda, mu (src/meshgen.f90:196-200)
static inline void set_mat(eqdm_arrays *A, int64_t e, int64_t ne_total, double vp, double vs, double rho)
{
    double *m = A->mat + (e - 1);  // mat(ne, 5), column stride ne
    m[0 * ne_total] = vp;
    m[1 * ne_total] = vs;
    m[2 * ne_total] = rho;
    m[4 * ne_total] = vs * vs * rho;
    m[3 * ne_total] = vp * vp * rho - 2.0 * m[4 * ne_total];
}

// Node loop of src/meshgen.f90:64-107 (topology only) or, with arrays, the
// full fill.  Returns counts.
static Counts topology(const eqdm_params &p, const Grid &G, eqdm_arrays *A, int64_t ne_total)
{
    Counts c;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int64_t nreg = (int64_t)nx * ny * nz;
    std::vector<int32_t> plane1((size_t)(ny + 2) * (nz + 1), 0), plane2((size_t)(ny + 2) * (nz + 1), 0);
    auto P1 = [&](int iy, int iz) -> int32_t & { return plane1[(size_t)iz * (ny + 2) + iy]; };
    auto P2 = [&](int iy, int iz) -> int32_t & { return plane2[(size_t)iz * (ny + 2) + iy]; };
    // y coordinates of the plane nodes (shifted by a rough fault surface)
    std::vector<double> ypl1((size_t)(ny + 2) * (nz + 1), 0.0), ypl2((size_t)(ny + 2) * (nz + 1), 0.0);
    auto Y1 = [&](int iy, int iz) -> double & { return ypl1[(size_t)iz * (ny + 2) + iy]; };
    auto Y2 = [&](int iy, int iz) -> double & { return ypl2[(size_t)iz * (ny + 2) + iy]; };
    int64_t node = 0, elem = 0, eq = 0, tag = 0, sdof = 0, nf = 0, non = 0, noff = 0;
    std::unordered_map<int32_t, int32_t> slave_to_master;
    std::vector<char> found(p.n_off, 0);
    std::vector<int32_t> stnode;  // OffFaultStNodeIdIndex(2,:)
    const double pi = 4.0 * std::atan(1.0);
    const double stk = p.fstrike * pi / 180.0, dip = (p.C_degen > 3.0 ? p.C_degen : 90.0) * pi / 180.0;
    // fault area bookkeeping (fltrc of src/meshgen.f90:770-775)
    int ixfi = 0, izfi = 0, ifs = 0, ifd = 0;
    std::vector<int32_t> frc_m, frc_k;  // [ifd][ifs] master node, pair index
    const int nxf = nx + 1, nzf = nz + 1;
    if (A) {
        frc_m.assign((size_t)nxf * nzf, 0);
        frc_k.assign((size_t)nxf * nzf, 0);
    }
    for (int ix = 1; ix <= nx; ix++) {
        for (int iz = 1; iz <= nz; iz++) {
            for (int iy = 1; iy <= ny; iy++) {
                const double x = G.x[ix], y = G.y[iy], z = G.z[iz];
                double yc = y, pfx = 0.0, pfz = 0.0;
                if (p.insertFaultType > 0) rough_shift(p, G.ymin, G.ymax, x, y, z, &yc, &pfx, &pfz);
                node++;
                P2(iy, iz) = (int32_t)node;
                Y2(iy, iz) = yc;
                const int ndof = is_pml_node(G, x, y, z) ? 12 : 3;
                const bool fixed = is_fixed(p, G, x, y, z);
                if (A) {
                    A->meshCoor[(node - 1) * 3] = x;
                    A->meshCoor[(node - 1) * 3 + 1] = yc;
                    A->meshCoor[(node - 1) * 3 + 2] = z;
                    A->eqNumStartIndexLoc[node - 1] = (int32_t)tag;
                    A->numOfDofPerNodeArr[node - 1] = ndof;
                }
                for (int d = 0; d < ndof; d++) {
                    tag++;
                    int32_t v = -1;
                    if (!fixed) v = (int32_t)(++eq);
                    if (A) A->eqNumIndexArr[tag - 1] = v;
                }
                // off-fault stations
                if (p.n_off > 0) {
                    const int st = match_station(p, G, ix, iy, x, y, z, found);
                    if (st >= 0) {
                        noff++;
                        stnode.push_back((int32_t)node);
                    }
                }
                // split node (createMasterNode, src/meshgen.f90:805-920)
                if (on_fault(p, x, y, z)) {
                    nf++;
                    const int64_t ms = nreg + nf;
                    P2(ny + 1, iz) = (int32_t)ms;
                    slave_to_master[(int32_t)node] = (int32_t)ms;
                    if (A) {
                        A->nsmp[(nf - 1) * 2] = (int32_t)node;
                        A->nsmp[(nf - 1) * 2 + 1] = (int32_t)ms;
                        A->eqNumStartIndexLoc[ms - 1] = (int32_t)tag;
                        A->numOfDofPerNodeArr[ms - 1] = 3;
                        A->meshCoor[(ms - 1) * 3] = x;
                        A->meshCoor[(ms - 1) * 3 + 1] = yc;
                        A->meshCoor[(ms - 1) * 3 + 2] = z;
                    }
                    for (int d = 0; d < 3; d++) {
                        eq++;
                        tag++;
                        if (A) A->eqNumIndexArr[tag - 1] = (int32_t)eq;
                    }
                    for (int i = 0; i < p.n_on; i++) {
                        if (std::fabs(x - p.st_on[i * 2]) < p.tol && std::fabs(z - p.st_on[i * 2 + 1]) < p.tol) {
                            non++;
                            if (A) {
                                A->anonfs[(non - 1) * 3] = (int32_t)nf;
                                A->anonfs[(non - 1) * 3 + 1] = i + 1;
                                A->anonfs[(non - 1) * 3 + 2] = 1;
                            }
                            break;
                        }
                    }
                    if (A) {
                        double *un = A->un + (nf - 1) * 3, *us = A->us + (nf - 1) * 3, *ud = A->ud + (nf - 1) * 3;
                        if (p.insertFaultType > 0) {
                            // normals from the local surface slopes, src/meshgen.f90:874-889
                            const double nrm = std::sqrt(pfx * pfx + 1.0 + pfz * pfz), sn = std::sqrt(1.0 + pfx * pfx);
                            un[0] = -pfx / nrm;
                            un[1] = 1.0 / nrm;
                            un[2] = -pfz / nrm;
                            us[0] = 1.0 / sn;
                            us[1] = pfx / sn;
                            us[2] = 0.0;
                            ud[0] = us[1] * un[2] - us[2] * un[1];
                            ud[1] = us[2] * un[0] - us[0] * un[2];
                            ud[2] = us[0] * un[1] - us[1] * un[0];
                        } else {
                            un[0] = std::cos(stk) * std::sin(dip);
                            un[1] = -std::sin(stk) * std::sin(dip);
                            un[2] = std::cos(dip);
                            us[0] = -std::sin(stk);
                            us[1] = -std::cos(stk);
                            us[2] = 0.0;
                            ud[0] = std::cos(stk) * std::cos(dip);
                            ud[1] = std::sin(stk) * std::cos(dip);
                            ud[2] = std::sin(dip);
                        }
                    }
                    if (ixfi == 0) ixfi = ix;
                    if (izfi == 0) izfi = iz;
                    ifs = ix - ixfi + 1;
                    ifd = iz - izfi + 1;
                    if (A) {
                        frc_m[(size_t)ifd * nxf + ifs] = (int32_t)ms;
                        frc_k[(size_t)ifd * nxf + ifs] = (int32_t)nf;
                    }
                }
                // element (createElement + setElementMaterial + replaceSlaveWithMasterNode)
                if (ix >= 2 && iy >= 2 && iz >= 2) {
                    // createElement, src/meshgen.f90:700-741
                    elem++;
                    const int32_t base[8] = {P1(iy - 1, iz - 1), P2(iy - 1, iz - 1), P2(iy, iz - 1), P1(iy, iz - 1),
                                             P1(iy - 1, iz),     P2(iy - 1, iz),     P2(iy, iz),     P1(iy, iz)};
                    // centroid from the nodes' coordinates (all nodes of the element exist already)
                    const int ixs[8] = {ix - 1, ix, ix, ix - 1, ix - 1, ix, ix, ix - 1};
                    const double ys[8] = {Y1(iy - 1, iz - 1), Y2(iy - 1, iz - 1), Y2(iy, iz - 1), Y1(iy, iz - 1),
                                          Y1(iy - 1, iz),     Y2(iy - 1, iz),     Y2(iy, iz),     Y1(iy, iz)};
                    const int izs[8] = {iz - 1, iz - 1, iz - 1, iz - 1, iz, iz, iz, iz};
                    double xc[3] = {0.0, 0.0, 0.0};
                    for (int a = 0; a < 8; a++) {
                        xc[0] = xc[0] + G.x[ixs[a]];
                        xc[1] = xc[1] + ys[a];
                        xc[2] = xc[2] + G.z[izs[a]];
                    }
                    for (int j = 0; j < 3; j++) xc[j] = xc[j] / 8.0;
                    const bool pml = xc[0] > G.PMLb[0] || xc[0] < G.PMLb[1] || xc[1] > G.PMLb[2] ||
                                     xc[1] < G.PMLb[3] || xc[2] < G.PMLb[4];
                    const int64_t e1 = elem;
                    int etype = pml ? 2 : 1;
                    if (A) {
                        A->stressCompIndexArr[e1 - 1] = (int32_t)sdof;
                        // setElementMaterial, src/meshgen.f90:158-204
                        double vp = 0, vs = 0, rho = 0;
                        if (p.nmat == 1 && p.n2mat == 3) {
                            vp = p.material[0]; vs = p.material[1]; rho = p.material[2];
                        } else if (p.nmat > 1 && p.n2mat == 4) {
                            const double dep = std::fabs(xc[2]);
                            if (dep < p.material[0]) {
                                vp = p.material[1]; vs = p.material[2]; rho = p.material[3];
                            } else {
                                for (int i = 1; i < p.nmat; i++)
                                    if (dep < p.material[i * 4] && dep >= p.material[(i - 1) * 4]) {
                                        vp = p.material[i * 4 + 1]; vs = p.material[i * 4 + 2]; rho = p.material[i * 4 + 3];
                                    }
                            }
                        }
                        set_mat(A, e1, ne_total, vp, vs, rho);
                    }
                    sdof += pml ? 21 : 12;
                    int32_t nd[8];
                    for (int a = 0; a < 8; a++) nd[a] = base[a];
                    int64_t last = e1;
                    if (p.C_degen > 3.0) {
                        // wedge, src/library_degeneration.f90:3-55: a hexahedron the
                        // dipping plane cuts along its diagonal becomes two wedges
                        const double tg = std::tan(p.C_degen / 180.0 * kPi);
                        const double d = std::fabs(xc[1] * tg + xc[2]) / std::sqrt(1.0 + tg * tg);
                        if (xc[0] > p.fxmin && xc[0] < p.fxmax && xc[1] > p.fymin && xc[1] < p.fymax &&
                            xc[2] > p.fzmin && d < p.tol) {
                            static const int below[8] = {5, 1, 4, 4, 6, 2, 3, 3}, above[8] = {4, 8, 5, 5, 3, 7, 6, 6};
                            if (A) {
                                A->elemTypeArr[e1 - 1] = 11;
                                for (int a = 0; a < 8; a++) A->nodeElemIdRelation[(e1 - 1) * 8 + a] = base[below[a] - 1];
                            }
                            elem++;
                            last = elem;
                            etype = 12;
                            for (int a = 0; a < 8; a++) nd[a] = base[above[a] - 1];
                            if (A) {
                                A->stressCompIndexArr[last - 1] = (int32_t)sdof;
                                // the upper wedge takes the first material row as given
                                set_mat(A, last, ne_total, p.material[0], p.material[1], p.material[2]);
                            }
                            sdof += 12;
                        } else if (etype == 1) {
                            // an element touching the plane with its nodes 1 or 2 sits on its upper side
                            const double y1 = ys[0], y2 = ys[1];
                            if (on_fault(p, G.x[ix - 1], y1, G.z[iz - 1]) || on_fault(p, G.x[ix], y2, G.z[iz - 1]))
                                etype = 13;
                        }
                    }
                    if (A) {
                        A->elemTypeArr[last - 1] = etype;
                        // replaceSlaveWithMasterNode, src/meshgen.f90:743-766
                        if ((etype == 1 && x > p.fxmin - p.tol && x < p.fxmax + p.dx + p.tol && z > p.fzmin - p.tol &&
                             y > 0.0 && std::fabs(y - p.dy) < p.tol) ||
                            etype == 12 || etype == 13) {
                            for (int a = 0; a < 8; a++) {
                                auto it = slave_to_master.find(nd[a]);
                                if (it != slave_to_master.end()) nd[a] = it->second;
                            }
                        }
                        for (int a = 0; a < 8; a++) A->nodeElemIdRelation[(last - 1) * 8 + a] = nd[a];
                        if (p.C_elastic == 0) {
                            // setPlasticStress, src/meshgen.f90:976-996
                            const double depth = -0.5 * (G.z[iz] + G.z[iz - 1]) + 7.3215;
                            const int et = etype == 2 ? 1 : 0;
                            const double ang = p.str1ToFaultAngle * pi / 180.0;
                            const double strVert = -(p.roumax - p.rhow * (p.gamar + 1.0)) * depth * p.grav;
                            const double devStr = std::fabs(strVert) * p.devStrToStrVertRatio;
                            double *sa = A->stressArr + A->stressCompIndexArr[last - 1];
                            sa[3 + 15 * et - 1] = strVert;
                            sa[1 + 15 * et - 1] = strVert - devStr * std::cos(2.0 * ang);
                            sa[2 + 15 * et - 1] = strVert + devStr * std::cos(2.0 * ang);
                            sa[6 + 15 * et - 1] = devStr * std::sin(2.0 * ang);
                        }
                    }
                }
            }
        }
        plane1 = plane2;
        ypl1 = ypl2;
    }
    c.nn = node + nf;
    c.ne = elem;
    c.neq = eq;
    c.size_eqidx = tag;
    c.size_stress = sdof;
    c.nf = nf;
    c.non = non;
    c.noff = noff;
    if (A) {
        // fault areas, src/meshgen.f90:114-152
        for (int i = 2; i <= ifd; i++)
            for (int j = 2; j <= ifs; j++) {
                const int32_t n1 = frc_m[(size_t)i * nxf + j], n2 = frc_m[(size_t)i * nxf + j - 1],
                              n3 = frc_m[(size_t)(i - 1) * nxf + j - 1], n4 = frc_m[(size_t)(i - 1) * nxf + j];
                const int32_t m1 = frc_k[(size_t)i * nxf + j], m2 = frc_k[(size_t)i * nxf + j - 1],
                              m3 = frc_k[(size_t)(i - 1) * nxf + j - 1], m4 = frc_k[(size_t)(i - 1) * nxf + j];
                auto X = [&](int32_t n, int k) { return A->meshCoor[(int64_t)(n - 1) * 3 + k]; };
                auto dist = [&](int32_t a, int32_t b) {
                    return std::sqrt((X(b, 0) - X(a, 0)) * (X(b, 0) - X(a, 0)) + (X(b, 1) - X(a, 1)) * (X(b, 1) - X(a, 1)) +
                                     (X(b, 2) - X(a, 2)) * (X(b, 2) - X(a, 2)));
                };
                const double aa1 = dist(n1, n2), bb1 = dist(n2, n3), cc1 = dist(n3, n4), dd1 = dist(n4, n1);
                const double p1 = dist(n2, n4), q1 = dist(n1, n3);
                const double t = bb1 * bb1 + dd1 * dd1 - aa1 * aa1 - cc1 * cc1;
                double area = 0.25 * std::sqrt(4 * p1 * p1 * q1 * q1 - t * t);
                area = 0.25 * area;
                A->arn[m1 - 1] = A->arn[m1 - 1] + area;
                A->arn[m2 - 1] = A->arn[m2 - 1] + area;
                A->arn[m3 - 1] = A->arn[m3 - 1] + area;
                A->arn[m4 - 1] = A->arn[m4 - 1] + area;
            }
        // idhist, src/eqdyna3d.f90:170-188
        int64_t row = 0;
        for (int64_t s = 0; s < noff; s++)
            for (int d = 1; d <= 3; d++)
                for (int k = 1; k <= 2; k++) {
                    A->idhist[row * 3] = stnode[s];
                    A->idhist[row * 3 + 1] = d;
                    A->idhist[row * 3 + 2] = k;
                    row++;
                }
    }
    return c;
}

}  // namespace

extern "C" {

int eqdm_size(const eqdm_params *p, eqdm_sizes *s)
{
    if (!p || !s) return -1;
    Grid G = make_grid(*p);
    Counts c = topology(*p, G, nullptr, 0);
    s->nx = G.nx;
    s->ny = G.ny;
    s->nz = G.nz;
    s->nn = c.nn;
    s->ne = c.ne;
    s->neq = c.neq;
    s->size_eqidx = c.size_eqidx;
    s->size_stress = c.size_stress;
    s->nf = c.nf;
    s->non = c.non;
    s->noff = c.noff;
    std::memcpy(s->PMLb, G.PMLb, sizeof(G.PMLb));
    s->xmin = G.xmin; s->xmax = G.xmax; s->ymin = G.ymin; s->ymax = G.ymax; s->zmin = G.zmin; s->zmax = G.zmax;
    return 0;
}

// Arrays must be zero-initialised by the caller and sized by eqdm_size.
// Returns 0, -2 on a non-positive Jacobian (src/calcGlobalShapeFunc.f90:52-57),
// -3 on an element without material (src/eqdyna3d.f90:214-224).
int eqdm_fill(const eqdm_params *p, const eqdm_sizes *s, eqdm_arrays *A, int32_t *bad_elem)
{
    if (!p || !s || !A) return -1;
    Grid G = make_grid(*p);
    const int64_t ne = s->ne;
    topology(*p, G, A, ne);
    // checkMeshMaterial, src/eqdyna3d.f90:214-224
    for (int64_t e = 0; e < ne; e++) {
        if (A->mat[e] == 0.0 || A->mat[ne + e] == 0.0 || A->mat[2 * ne + e] == 0.0) {
            if (bad_elem) *bad_elem = (int32_t)(e + 1);
            return -3;
        }
    }
    const double w = 8.0;
    int32_t bad = 0;
#pragma omp parallel for schedule(static) num_threads(p->nthreads > 0 ? p->nthreads : 1)
    for (int64_t e = 0; e < ne; e++) {
        double xl[3][8];
        for (int a = 0; a < 8; a++) {
            const int64_t nd = A->nodeElemIdRelation[e * 8 + a] - 1;
            for (int j = 0; j < 3; j++) xl[j][a] = A->meshCoor[nd * 3 + j];
        }
        double det;
        const bool wedge = A->elemTypeArr[e] == 11 || A->elemTypeArr[e] == 12;
        if (!eqd::elem_precompute(xl, wedge, A->mat[2 * ne + e], A->mat[3 * ne + e], A->mat[4 * ne + e], w,
                             A->eleshp + e * 24, &det, A->elemass + e * 24, A->ss + e * 6, A->phi + e * 32)) {
#pragma omp critical
            bad = (int32_t)(e + 1);
            continue;
        }
        A->eledet[e] = det;
    }
    if (bad) {
        if (bad_elem) *bad_elem = bad;
        return -2;
    }
    // nodal masses in element order, src/assembleGlobalMass.f90:283-326
    for (int64_t e = 0; e < ne; e++) {
        for (int a = 0; a < 8; a++) {
            const int64_t nd = A->nodeElemIdRelation[e * 8 + a] - 1;
            const int64_t b = A->eqNumStartIndexLoc[nd];
            const double *em = A->elemass + e * 24 + a * 3;
            if (A->numOfDofPerNodeArr[nd] == 12) {
                for (int k = 0; k < 3; k++) {
                    for (int j = 3 * k + 1; j <= 3 * k + 3; j++) {
                        const int32_t q = A->eqNumIndexArr[b + j - 1];
                        if (q > 0) A->nodalMassArr[q - 1] = A->nodalMassArr[q - 1] + em[k];
                    }
                    const int32_t q = A->eqNumIndexArr[b + k + 9];
                    if (q > 0) A->nodalMassArr[q - 1] = A->nodalMassArr[q - 1] + em[k];
                }
            } else {
                for (int j = 1; j <= 3; j++) {
                    const int32_t q = A->eqNumIndexArr[b + j - 1];
                    A->nodalMassArr[q - 1] = A->nodalMassArr[q - 1] + em[j - 1];
                }
            }
            A->fnms[nd] = A->fnms[nd] + em[0];
        }
    }
    return 0;
}

}  // extern "C"
