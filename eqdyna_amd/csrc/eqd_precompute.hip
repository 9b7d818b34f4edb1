// Device-side one-time precompute: the element part of the reference host's
// assembleGlobalMass (src/assembleGlobalMass.f90:283-406 with
// src/calcGlobalShapeFunc.f90:1-75 and src/library.f90:60-93) -- per element
// dN/dx, Jacobian, lumped mass, SS and Phi (eqd_precompute.h, one thread per
// element), then the nodal masses.  The reference adds every element's lumped
// mass to its nodes' equations in ascending element order; here each node
// sums its own (element, local node) entries in that order, from a
// node-to-element table built on the device (counts by atomics, positions by
// a prefix sum, then each node's few entries sorted), so the sums are the
// reference's, bit for bit, without atomics on the masses.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/eqdyna_amd.h"
#include "eqd_precompute.h"

namespace eqd {
namespace {

__global__ void __launch_bounds__(128) elem_precompute_kernel(const double *__restrict__ coor,
                                                              const int32_t *__restrict__ ien,
                                                              const int32_t *__restrict__ etype,
                                                              const double *__restrict__ mat, int64_t ne, int64_t nn,
                                                              double w, double *shp, double *det, double *emass,
                                                              double *ss, double *phi, int32_t *bad)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double xl[3][8];
    for (int a = 0; a < 8; a++) {
        int64_t nd = (int64_t)ien[e * 8 + a] - 1;
        if (nd < 0 || nd >= nn) {  // checked on the host; kept in range regardless
            atomicMin(bad, (int32_t)(e + 1));
            return;
        }
        for (int j = 0; j < 3; j++) xl[j][a] = coor[nd * 3 + j];
    }
    const bool wedge = etype[e] == 11 || etype[e] == 12;
    double d = 0.0;
    if (!elem_precompute(xl, wedge, mat[2 * ne + e], mat[3 * ne + e], mat[4 * ne + e], w, shp + e * 24, &d,
                         emass + e * 24, ss + e * 6, phi + e * 32)) {
        atomicMin(bad, (int32_t)(e + 1));
        return;
    }
    det[e] = d;
}

// entries per node: one per (element, local node), wedges repeat nodes
__global__ void count_kernel(const int32_t *__restrict__ ien, int64_t n8, int32_t *cnt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    atomicAdd(&cnt[ien[i] - 1], 1);
}

__global__ void fill_kernel(const int32_t *__restrict__ ien, int64_t n8, const int32_t *__restrict__ rowptr,
                            int32_t *fillc, int64_t *ent)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    const int nd = ien[i] - 1;
    const int pos = rowptr[nd] + atomicAdd(&fillc[nd], 1);
    ent[pos] = i;  // e * 8 + a: ascending entry order is the reference's order
}

// each node's entries into ascending (element, local node) order, then its
// masses summed in that order (src/assembleGlobalMass.f90:283-326)
__global__ void node_mass_kernel(const int32_t *__restrict__ rowptr, int64_t *ent, const double *__restrict__ emass,
                                 const int32_t *__restrict__ ndof, const int32_t *__restrict__ esl,
                                 const int32_t *__restrict__ eqi, int64_t nn, double *nodal_mass, double *fnms)
{
    const int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= nn) return;
    const int beg = rowptr[nd], end = rowptr[nd + 1];
    for (int i = beg + 1; i < end; i++) {  // insertion sort: a node has at most 16 entries
        const int64_t v = ent[i];
        int j = i - 1;
        while (j >= beg && ent[j] > v) {
            ent[j + 1] = ent[j];
            j--;
        }
        ent[j + 1] = v;
    }
    const int64_t b = esl[nd];
    const bool pml = ndof[nd] == 12;
    double f = fnms[nd];
    for (int i = beg; i < end; i++) {
        const double *em = emass + ent[i] * 3;  // (e * 8 + a) * 3: element e, local node a
        if (pml) {
            for (int k = 0; k < 3; k++) {
                for (int j = 3 * k + 1; j <= 3 * k + 3; j++) {
                    const int32_t q = eqi[b + j - 1];
                    if (q > 0) nodal_mass[q - 1] = nodal_mass[q - 1] + em[k];
                }
                const int32_t q = eqi[b + k + 9];
                if (q > 0) nodal_mass[q - 1] = nodal_mass[q - 1] + em[k];
            }
        } else {
            for (int j = 1; j <= 3; j++) {
                const int32_t q = eqi[b + j - 1];
                if (q > 0) nodal_mass[q - 1] = nodal_mass[q - 1] + em[j - 1];
            }
        }
        f = f + em[0];
    }
    fnms[nd] = f;
}

struct DevBuf {
    std::vector<void *> p;
    ~DevBuf()
    {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T>
    hipError_t get(T **out, size_t count)
    {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, sizeof(T) * (count ? count : 1));
        if (e == hipSuccess) p.push_back(q);
        *out = (T *)q;
        return e;
    }
};

}  // namespace
}  // namespace eqd

using namespace eqd;

#define PCHK(x)                                      \
    do {                                             \
        if ((x) != hipSuccess) return EQD_EHIP;      \
    } while (0)

extern "C" int eqd_precompute(const eqd_precompute_args *A, int32_t *bad_elem, double *device_ms)
{
    if (!A || A->nn <= 0 || A->ne <= 0 || A->neq <= 0 || A->size_eqidx <= 0) return EQD_EINVAL;
    if (!A->meshCoor || !A->nodeElemIdRelation || !A->elemTypeArr || !A->mat || !A->numOfDofPerNodeArr ||
        !A->eqNumStartIndexLoc || !A->eqNumIndexArr || !A->eleshp || !A->eledet || !A->elemass || !A->ss ||
        !A->phi || !A->nodalMassArr || !A->fnms)
        return EQD_EINVAL;
    const int64_t nn = A->nn, ne = A->ne, neq = A->neq;
    if (bad_elem) *bad_elem = 0;
    // the host checks every index the kernels use
    for (int64_t i = 0; i < 8 * ne; i++)
        if (A->nodeElemIdRelation[i] < 1 || A->nodeElemIdRelation[i] > nn) {
            if (bad_elem) *bad_elem = (int32_t)(i / 8 + 1);
            return EQD_EINVAL;
        }
    for (int64_t n = 0; n < nn; n++) {
        const int w = A->numOfDofPerNodeArr[n] == 12 ? 12 : 3;
        const int64_t b = A->eqNumStartIndexLoc[n];
        if (b < 0 || b + w > A->size_eqidx) return EQD_EINVAL;
        for (int j = 0; j < w; j++)
            if (A->eqNumIndexArr[b + j] > neq) return EQD_EINVAL;
    }
    PCHK(hipSetDevice(A->device));
    hipStream_t st;
    PCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    DevBuf B;
    double *coor, *mat, *shp, *det, *em, *ss, *phi, *nm, *fn;
    int32_t *ien, *et, *ndof, *esl, *eqi, *cnt, *rowptr, *fillc, *bad;
    int64_t *ent;
    PCHK(B.get(&coor, 3 * nn));
    PCHK(B.get(&mat, 5 * ne));
    PCHK(B.get(&ien, 8 * ne));
    PCHK(B.get(&et, ne));
    PCHK(B.get(&shp, 24 * ne));
    PCHK(B.get(&det, ne));
    PCHK(B.get(&em, 24 * ne));
    PCHK(B.get(&ss, 6 * ne));
    PCHK(B.get(&phi, 32 * ne));
    PCHK(B.get(&ndof, nn));
    PCHK(B.get(&esl, nn));
    PCHK(B.get(&eqi, A->size_eqidx));
    PCHK(B.get(&nm, neq));
    PCHK(B.get(&fn, nn));
    PCHK(B.get(&cnt, nn + 1));
    PCHK(B.get(&rowptr, nn + 1));
    PCHK(B.get(&fillc, nn));
    PCHK(B.get(&ent, 8 * ne));
    PCHK(B.get(&bad, 1));
    auto h2d = [&](void *d, const void *h, size_t b) { return hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, st); };
    PCHK(h2d(coor, A->meshCoor, sizeof(double) * 3 * nn));
    PCHK(h2d(mat, A->mat, sizeof(double) * 5 * ne));
    PCHK(h2d(ien, A->nodeElemIdRelation, sizeof(int32_t) * 8 * ne));
    PCHK(h2d(et, A->elemTypeArr, sizeof(int32_t) * ne));
    PCHK(h2d(ndof, A->numOfDofPerNodeArr, sizeof(int32_t) * nn));
    PCHK(h2d(esl, A->eqNumStartIndexLoc, sizeof(int32_t) * nn));
    PCHK(h2d(eqi, A->eqNumIndexArr, sizeof(int32_t) * A->size_eqidx));
    PCHK(h2d(nm, A->nodalMassArr, sizeof(double) * neq));
    PCHK(h2d(fn, A->fnms, sizeof(double) * nn));
    PCHK(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (nn + 1), st));
    PCHK(hipMemsetAsync(fillc, 0, sizeof(int32_t) * nn, st));
    const int32_t big = 0x7fffffff;
    PCHK(h2d(bad, &big, sizeof(int32_t)));
    hipEvent_t ev0, ev1;
    PCHK(hipEventCreate(&ev0));
    PCHK(hipEventCreate(&ev1));
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{ev0, ev1};
    PCHK(hipEventRecord(ev0, st));
    const unsigned ge = (unsigned)((ne + 127) / 128);
    elem_precompute_kernel<<<ge, 128, 0, st>>>(coor, ien, et, mat, ne, nn, 8.0, shp, det, em, ss, phi, bad);
    PCHK(hipGetLastError());
    const int64_t n8 = 8 * ne;
    const unsigned g8 = (unsigned)((n8 + 255) / 256), gn = (unsigned)((nn + 255) / 256);
    count_kernel<<<g8, 256, 0, st>>>(ien, n8, cnt);
    PCHK(hipGetLastError());
    size_t tmp_bytes = 0;
    PCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, cnt, rowptr, (int)(nn + 1), st));
    void *tmp;
    PCHK(B.get((unsigned char **)&tmp, tmp_bytes));
    PCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, rowptr, (int)(nn + 1), st));
    fill_kernel<<<g8, 256, 0, st>>>(ien, n8, rowptr, fillc, ent);
    PCHK(hipGetLastError());
    int32_t hbad = 0;
    PCHK(hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PCHK(hipStreamSynchronize(st));
    if (hbad != big) {
        if (bad_elem) *bad_elem = hbad;
        return EQD_EINVAL;
    }
    node_mass_kernel<<<gn, 256, 0, st>>>(rowptr, ent, em, ndof, esl, eqi, nn, nm, fn);
    PCHK(hipGetLastError());
    PCHK(hipEventRecord(ev1, st));
    auto d2h = [&](void *h, const void *d, size_t b) { return hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, st); };
    PCHK(d2h(A->eleshp, shp, sizeof(double) * 24 * ne));
    PCHK(d2h(A->eledet, det, sizeof(double) * ne));
    PCHK(d2h(A->elemass, em, sizeof(double) * 24 * ne));
    PCHK(d2h(A->ss, ss, sizeof(double) * 6 * ne));
    PCHK(d2h(A->phi, phi, sizeof(double) * 32 * ne));
    PCHK(d2h(A->nodalMassArr, nm, sizeof(double) * neq));
    PCHK(d2h(A->fnms, fn, sizeof(double) * nn));
    PCHK(hipStreamSynchronize(st));
    if (device_ms) {
        float ms = 0.0f;
        PCHK(hipEventElapsedTime(&ms, ev0, ev1));
        *device_ms = ms;
    }
    return EQD_OK;
}
