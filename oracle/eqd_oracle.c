/* Test infrastructure, not product code.
 *
 * CPU FP64 restatement of the EQdyna 5.3.3 explicit time step.  Every
 * function follows one reference routine, cited by file:line, and keeps its
 * floating-point operation order (Fortran evaluates a+b+c as (a+b)+c and
 * -a*b*c as -((a*b)*c)), so that with FMA contraction off (-ffp-contract=off)
 * it agrees with the reference build in oracle/_ref to the last bit wherever
 * the math library agrees.  Arrays are the reference's own (column-major,
 * 1-based ids); the macros below do the index arithmetic.
 */
#include "eqd_oracle.h"

#include <math.h>
#include <stddef.h>
#include <string.h>

#define MC(k, n) st->meshCoor[(size_t)((n) - 1) * 3 + (k) - 1]
#define EQI(i) st->eqNumIndexArr[(i) - 1]
#define ESL(n) st->eqNumStartIndexLoc[(n) - 1]
#define NDOF(n) st->numOfDofPerNodeArr[(n) - 1]
#define IEN(a, e) st->ien[(size_t)((e) - 1) * 8 + (a) - 1]
#define MAT(e, k) st->mat[(size_t)((k) - 1) * s->ne + (e) - 1]
#define SHP(k, a, e) st->eleshp[(size_t)((e) - 1) * 24 + ((a) - 1) * 3 + (k) - 1]
#define EMASS(i, e) st->elemass[(size_t)((e) - 1) * 24 + (i) - 1]
#define SS(k, e) st->ss[(size_t)((e) - 1) * 6 + (k) - 1]
#define PHI(a, m, e) st->phi[(size_t)((e) - 1) * 32 + ((m) - 1) * 8 + (a) - 1]
#define VEL(k, n) st->velArr[(size_t)((n) - 1) * 3 + (k) - 1]
#define DISP(k, n) st->dispArr[(size_t)((n) - 1) * 3 + (k) - 1]
#define V1(i) st->v1[(i) - 1]
#define FORCE(i) st->nodalForceArr[(i) - 1]
#define FRIC(k, i) st->fric[(size_t)((i) - 1) * 100 + (k) - 1]
#define NSMP(j, i) st->nsmp[(size_t)((i) - 1) * 2 + (j) - 1]
#define UN(k, i) st->un[(size_t)((i) - 1) * 3 + (k) - 1]
#define US(k, i) st->us[(size_t)((i) - 1) * 3 + (k) - 1]
#define UD(k, i) st->ud[(size_t)((i) - 1) * 3 + (k) - 1]

static double sq(double x) { return x * x; }

/* ------------------------------------------------------------------ comdampv
 * src/comdampv.f90:1-120.  Region 14 tests y >= xmax0 (not ymax0), as the
 * reference does. */
void eqdo_comdampv(const eqdo_scalars *s, double x2, double y, double z, double dv[9])
{
    double xmax0 = s->PMLb[0], xmin0 = s->PMLb[1], ymax0 = s->PMLb[2], ymin0 = s->PMLb[3];
    double zmin0 = s->PMLb[4], maxdx = s->PMLb[5], maxdy = s->PMLb[6], maxdz = s->PMLb[7];
    double damp[3];
    if (z <= zmin0)
        damp[2] = fabs(z - zmin0);
    else
        damp[2] = 0.0;
    if (x2 >= xmax0 && y >= ymax0) {
        damp[0] = fabs(x2 - xmax0); damp[1] = fabs(y - ymax0);
    } else if (x2 >= xmax0 && y <= ymin0) {
        damp[0] = fabs(x2 - xmax0); damp[1] = fabs(y - ymin0);
    } else if (x2 <= xmin0 && y <= ymin0) {
        damp[0] = fabs(x2 - xmin0); damp[1] = fabs(y - ymin0);
    } else if (x2 <= xmin0 && y >= xmax0) {
        damp[0] = fabs(x2 - xmin0); damp[1] = fabs(y - ymax0);
    } else if (x2 >= xmax0 && y > ymin0 && y < ymax0) {
        damp[0] = fabs(x2 - xmax0); damp[1] = 0.0;
    } else if (y <= ymin0 && x2 > xmin0 && x2 < xmax0) {
        damp[0] = 0.0; damp[1] = fabs(y - ymin0);
    } else if (x2 <= xmin0 && y > ymin0 && y < ymax0) {
        damp[0] = fabs(x2 - xmin0); damp[1] = 0.0;
    } else if (y >= ymax0 && x2 > xmin0 && x2 < xmax0) {
        damp[0] = 0.0; damp[1] = fabs(y - ymax0);
    } else {
        damp[0] = 0.0; damp[1] = 0.0;
    }
    for (int i = 0; i < 3; i++) {
        double delta = (double)s->nPML * (i == 0 ? maxdx : (i == 1 ? maxdy : maxdz));
        damp[i] = 3.0 * s->vmaxPML / 2.0 / delta * log(1.0 / s->R) * pow(damp[i] / delta, 2.0);
    }
    for (int i = 0; i < 9; i++) dv[i] = damp[i % 3];
}

/* -------------------------------------------------------------- velDispUpdate
 * src/driver.f90:89-155 */
static int vel_disp_update(const eqdo_scalars *s, eqdo_state *st)
{
    const double dt = s->dt;
    double dampv[9] = {0};
    for (int i = 1; i <= s->nn; i++) {
        if (NDOF(i) == 3) {
            for (int j = 1; j <= 3; j++) {
                int eq = EQI(ESL(i) + j);
                V1(eq) = V1(eq) + FORCE(eq) * dt;
                VEL(j, i) = V1(eq);
                DISP(j, i) = DISP(j, i) + V1(eq) * dt;
            }
        } else if (NDOF(i) == 12) {
            int eq = EQI(ESL(i) + 1);
            if (eq > 0) eqdo_comdampv(s, MC(1, i), MC(2, i), MC(3, i), dampv);
            for (int j = 1; j <= 9; j++) {
                eq = EQI(ESL(i) + j);
                if (eq > 0)
                    V1(eq) = (FORCE(eq) + V1(eq) * (1.0 / dt - dampv[j - 1] / 2.0)) / (1.0 / dt + dampv[j - 1] / 2.0);
            }
            for (int j = 10; j <= 12; j++) {
                eq = EQI(ESL(i) + j);
                if (eq > 0) V1(eq) = V1(eq) + FORCE(eq) * dt;
            }
            eq = EQI(ESL(i) + 1);
            if (eq > 0) {
                int b = ESL(i);
                VEL(1, i) = V1(EQI(b + 1)) + V1(EQI(b + 2)) + V1(EQI(b + 3)) + V1(EQI(b + 10));
                VEL(2, i) = V1(EQI(b + 4)) + V1(EQI(b + 5)) + V1(EQI(b + 6)) + V1(EQI(b + 11));
                VEL(3, i) = V1(EQI(b + 7)) + V1(EQI(b + 8)) + V1(EQI(b + 9)) + V1(EQI(b + 12));
                DISP(1, i) = DISP(1, i) + VEL(1, i) * dt;
                DISP(2, i) = DISP(2, i) + VEL(2, i) * dt;
                DISP(3, i) = DISP(3, i) + VEL(3, i) * dt;
            } else if (eq == -1) {
                for (int k = 1; k <= 3; k++) { VEL(k, i) = 0.0; DISP(k, i) = 0.0; }
            }
        }
        if (VEL(1, i) != VEL(1, i) || VEL(2, i) != VEL(2, i) || VEL(3, i) != VEL(3, i)) {
            st->err_node = i;
            st->err_nt = st->nt;
            return 1;
        }
    }
    return 0;
}

/* ---------------------------------------------------------- storeOffFaultStData
 * src/driver.f90:157-180 */
static void store_off_fault(const eqdo_scalars *s, eqdo_state *st)
{
    int nrow = s->noff * 6;
    if (nrow <= 0 || st->offFault == NULL) return;
    double *col = st->offFault + (size_t)(st->nt - 1) * (nrow + 1);
    col[0] = st->timeElapsed;
    for (int i = 1; i <= nrow; i++) {
        int node = st->idhist[(i - 1) * 3 + 0];
        int k = st->idhist[(i - 1) * 3 + 1];
        int q = st->idhist[(i - 1) * 3 + 2];
        if (q == 1)
            col[i] = DISP(k, node);
        else if (q == 2)
            col[i] = VEL(k, node);
        else if (q == 3)
            col[i] = FORCE(EQI(ESL(node) + k));
    }
}

/* ---------------------------------------------------------------- qconstant
 * src/qconstant.f90:3-34.  The coefficient literals are default REAL in the
 * reference, so each is rounded to float before use. */
void eqdo_qconstant(double Q, int32_t k, double pi, double *rtaok, double *rwk, double *c1)
{
    static const float taok[8] = {1.72333e-3f, 1.80701e-3f, 5.38887e-3f, 1.99322e-2f,
                                  8.49833e-2f, 4.09335e-1f, 2.05951f, 13.2629f};
    static const float alfk[8] = {1.66958e-2f, 3.81644e-2f, 9.84666e-3f, -1.36803e-2f,
                                  -2.85125e-2f, -5.37309e-2f, -6.65035e-2f, -1.33696e-1f};
    static const float betk[8] = {8.98758e-2f, 6.84635e-2f, 9.67052e-2f, 1.20172e-1f,
                                  1.30728e-1f, 1.38746e-1f, 1.40705e-1f, 2.14647e-1f};
    double kapa = (double)3.071f + (double)1.433f * pow(Q, (double)(-1.158f)) * log(Q / 5.0);
    kapa = kapa / (1.0 + (double)0.415f * Q);
    *rwk = kapa * (kapa * (double)alfk[k - 1] + (double)betk[k - 1]);
    *rtaok = (double)taok[k - 1];
    double ref = 2.0 * pi;
    double ak0 = 1.0 - *rwk * 8.0 / (1.0 + sq(*rtaok * ref));
    double bk0 = *rwk * 8.0 * ref * *rtaok / (1.0 + sq(*rtaok * ref));
    /* x**(-0.5) is lowered by flang to 1/sqrt(x) */
    double rs = 1.0 / sqrt(sq(ak0) + sq(bk0));
    double c = 0.5 * rs;
    c = c * (1.0 + ak0 * rs);
    *c1 = c;
}

/* --------------------------------------------------------------- calcElemKU
 * src/calcElemKU.f90:3-191 with calcB (src/calcB.f90:1-26) folded into the
 * index pattern: b(1,j1)=N1 b(5,j1)=N3 b(6,j1)=N2; b(2,j2)=N2 b(4,j2)=N3
 * b(6,j2)=N1; b(3,j3)=N3 b(4,j3)=N2 b(5,j3)=N1 (N = derivatives of node i). */
static void calc_elem_ku(const eqdo_scalars *s, const eqdo_state *st, int e, const double vl[24],
                         const double dl[24], double *stress, double elresf[24], double constk,
                         double porep, double *pstrmag, const double ex[24])
{
    double sr[6] = {0}, strain[6] = {0}, srt[6] = {0}, strdev[6], strtemp[6], pstrinc[6];
    double lam = MAT(e, 4), miu = MAT(e, 5);
    double cdiag = lam + 2 * miu;
    *pstrmag = 0.0;
    for (int i = 1; i <= 8; i++) {
        double n1 = SHP(1, i, e), n2 = SHP(2, i, e), n3 = SHP(3, i, e);
        int j1 = 3 * (i - 1), j2 = j1 + 1, j3 = j1 + 2;
        sr[0] = sr[0] + n1 * vl[j1];
        sr[1] = sr[1] + n2 * vl[j2];
        sr[2] = sr[2] + n3 * vl[j3];
        sr[3] = sr[3] + n3 * vl[j2] + n2 * vl[j3];
        sr[4] = sr[4] + n3 * vl[j1] + n1 * vl[j3];
        sr[5] = sr[5] + n2 * vl[j1] + n1 * vl[j2];
        strain[0] = strain[0] + n1 * dl[j1];
        strain[1] = strain[1] + n2 * dl[j2];
        strain[2] = strain[2] + n3 * dl[j3];
        strain[3] = strain[3] + n3 * dl[j2] + n2 * dl[j3];
        strain[4] = strain[4] + n3 * dl[j1] + n1 * dl[j3];
        strain[5] = strain[5] + n2 * dl[j1] + n1 * dl[j2];
    }
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) srt[i] = srt[i] + (i == j ? cdiag : lam) * sr[j];
    }
    for (int i = 3; i < 6; i++) srt[i] = miu * sr[i];
    if (s->C_Q == 0) {
        for (int i = 0; i < 6; i++) {
            stress[i] = stress[i] + srt[i] * s->dt;
            strdev[i] = stress[i];
        }
    } else if (s->C_Q == 1) {
        double xc[3] = {0, 0, 0}, Qs, Qp, taok, wkp, wks, cv, cs, anestr[6], anestr1[6];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 8; j++) xc[i] = xc[i] + ex[j * 3 + i];
        for (int i = 0; i < 3; i++) xc[i] = xc[i] / 8.0;
        if (xc[2] > -1000.0) { Qs = 10.0; Qp = 20.0; } else { Qs = 50.0; Qp = 100.0; }
        int ip = (int)((xc[0] - (s->PMLb[1] + s->dx / 2)) / s->dx + 1);
        int iq = (int)((xc[1] - (s->PMLb[3] + s->dx / 2)) / s->dx + 1);
        int ir = (int)((xc[2] - (s->PMLb[4] + s->dx / 2)) / s->dx + 1);
        int k = 1 + ip % 2 + 2 * (iq % 2) + 4 * (ir % 2);
        eqdo_qconstant(Qp, k, s->pi, &taok, &wkp, &cv);
        eqdo_qconstant(Qs, k, s->pi, &taok, &wks, &cs);
        wkp = wkp * 8.0;
        wks = wks * 8.0;
        double miuu = miu * cs;
        double Mu = (lam + 2 * miu) * cv;
        double vols = strain[0] + strain[1] + strain[2];
        for (int i = 0; i < 6; i++) anestr1[i] = stress[i + 6];
        for (int i = 0; i < 3; i++)
            anestr[i] = exp(-s->dt / taok) * anestr1[i] + (1 - exp(-s->dt / taok)) *
                        (2 * miuu * strain[i] * wks + (Mu * wkp - 2 * miuu * wks) * vols);
        for (int i = 3; i < 6; i++)
            anestr[i] = exp(-s->dt / taok) * anestr1[i] + (1 - exp(-s->dt / taok)) * (miuu * strain[i] * wks);
        for (int i = 0; i < 6; i++) stress[i + 6] = anestr[i];
        for (int i = 0; i < 3; i++)
            stress[i] = 2.0 * miuu * strain[i] + (Mu - 2.0 * miuu) * vols - 0.5 * (anestr[i] + anestr1[i]);
        for (int i = 3; i < 6; i++)
            stress[i] = 2.0 * miuu * strain[i] / 2.0 - 0.5 * (anestr[i] + anestr1[i]);
    }
    if (s->C_elastic == 0) {
        double strmea = (stress[0] + stress[1] + stress[2]) / 3.0;
        for (int i = 0; i < 3; i++) strdev[i] = stress[i] - strmea;
        /* strdev(4:6) keep the values set in the C_Q==0 branch, as in the reference */
        double taomax = 0.5 * (sq(strdev[0]) + sq(strdev[1]) + sq(strdev[2])) + sq(strdev[3]) +
                        sq(strdev[4]) + sq(strdev[5]);
        taomax = sqrt(taomax);
        double yield = s->ccosphi - s->sinphi * (strmea + porep);
        if (yield < 0.0) yield = 0.0;
        if (taomax > yield) {
            double rjust = yield / taomax + (1 - yield / taomax) * exp(-s->dt / s->tv);
            for (int i = 0; i < 6; i++) {
                stress[i] = strdev[i] * rjust;
                pstrinc[i] = (strdev[i] - stress[i]) / miu;
                if (i < 3) stress[i] = stress[i] + strmea;
            }
            double pstrmea = (pstrinc[0] + pstrinc[1] + pstrinc[2]) / 3.0;
            for (int i = 0; i < 6; i++) pstrinc[i] = pstrinc[i] - pstrmea;
            double pm = 0.5 * (sq(pstrinc[0]) + sq(pstrinc[1]) + sq(pstrinc[2])) + sq(pstrinc[3]) +
                        sq(pstrinc[4]) + sq(pstrinc[5]);
            *pstrmag = sqrt(pm);
        }
    }
    double temp = constk * s->w;
    for (int i = 0; i < 6; i++) strtemp[i] = temp * (stress[i] + s->rdampk * srt[i]);
    for (int i = 1; i <= 8; i++) {
        double n1 = SHP(1, i, e), n2 = SHP(2, i, e), n3 = SHP(3, i, e);
        int j1 = 3 * (i - 1);
        double w1 = n1 * strtemp[0] + n3 * strtemp[4] + n2 * strtemp[5];
        double w2 = n2 * strtemp[1] + n3 * strtemp[3] + n1 * strtemp[5];
        double w3 = n3 * strtemp[2] + n2 * strtemp[3] + n1 * strtemp[4];
        elresf[j1] = elresf[j1] + w1;
        elresf[j1 + 1] = elresf[j1 + 1] + w2;
        elresf[j1 + 2] = elresf[j1 + 2] + w3;
    }
}

/* ------------------------------------------------------------ calcPMLElemKU
 * src/assembleGlobalKU.f90:70-346 */
static void calc_pml_elem_ku(const eqdo_scalars *s, const eqdo_state *st, int e, const double va[24],
                             double f[96], double *sv, const double ex[24])
{
    double xmax2 = s->PMLb[0], xmin2 = s->PMLb[1], ymax2 = s->PMLb[2], ymin2 = s->PMLb[3];
    double zmin2 = s->PMLb[4];
    double vp = MAT(e, 1), lam = MAT(e, 4), miu = MAT(e, 5), det = st->eledet[e - 1];
    double xc[3] = {0, 0, 0}, damps[3];
    (void)vp;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 8; j++) xc[i] = xc[i] + ex[j * 3 + i];
    for (int i = 0; i < 3; i++) xc[i] = xc[i] / 8;
    if (xc[2] < zmin2)
        damps[2] = fabs(xc[2] - zmin2);
    else if (xc[2] > zmin2)
        damps[2] = 0.0;
    else
        damps[2] = 0.0; /* xc(3)==zmin2 leaves damps uninitialised in the reference; never met on a grid */
    if (xc[2] < zmin2 || xc[2] > zmin2) {
        if (xc[0] > xmax2 && xc[1] > ymax2) {
            damps[0] = fabs(xc[0] - xmax2); damps[1] = fabs(xc[1] - ymax2);
        } else if (xc[0] > xmax2 && xc[1] < ymin2) {
            damps[0] = fabs(xc[0] - xmax2); damps[1] = fabs(xc[1] - ymin2);
        } else if (xc[0] < xmin2 && xc[1] < ymin2) {
            damps[0] = fabs(xc[0] - xmin2); damps[1] = fabs(xc[1] - ymin2);
        } else if (xc[0] < xmin2 && xc[1] > xmax2) {
            damps[0] = fabs(xc[0] - xmin2); damps[1] = fabs(xc[1] - ymax2);
        } else if (xc[0] > xmax2 && xc[1] > ymin2 && xc[1] < ymax2) {
            damps[0] = fabs(xc[0] - xmax2); damps[1] = 0.0;
        } else if (xc[1] < ymin2 && xc[0] > xmin2 && xc[0] < xmax2) {
            damps[0] = 0.0; damps[1] = fabs(xc[1] - ymin2);
        } else if (xc[0] < xmin2 && xc[1] > ymin2 && xc[1] < ymax2) {
            damps[0] = fabs(xc[0] - xmin2); damps[1] = 0.0;
        } else if (xc[1] > ymax2 && xc[0] > xmin2 && xc[0] < xmax2) {
            damps[0] = 0.0; damps[1] = fabs(xc[1] - ymax2);
        } else {
            damps[0] = 0.0; damps[1] = 0.0;
        }
    } else {
        damps[0] = 0.0; damps[1] = 0.0;
    }
    for (int i = 0; i < 3; i++) {
        double delta = (double)s->nPML * s->PMLb[5 + i];
        damps[i] = 3 * s->vmaxPML / 2 / delta * log(1 / s->R) * pow(damps[i] / delta, (double)2.0f);
    }
    double sr[6] = {0}, srt[6] = {0};
    for (int i = 1; i <= 8; i++) {
        double n1 = SHP(1, i, e), n2 = SHP(2, i, e), n3 = SHP(3, i, e);
        int j1 = 3 * (i - 1), j2 = j1 + 1, j3 = j1 + 2;
        sr[0] = sr[0] + n1 * va[j1];
        sr[1] = sr[1] + n2 * va[j2];
        sr[2] = sr[2] + n3 * va[j3];
        sr[3] = sr[3] + n3 * va[j2] + n2 * va[j3];
        sr[4] = sr[4] + n3 * va[j1] + n1 * va[j3];
        sr[5] = sr[5] + n2 * va[j1] + n1 * va[j2];
    }
    double cdiag = lam + 2.0 * miu;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) srt[i] = srt[i] + (i == j ? cdiag : lam) * sr[j];
    for (int i = 3; i < 6; i++) srt[i] = miu * sr[i];
    double Dx_vx = 0, Dy_vy = 0, Dz_vz = 0, Dx_vy = 0, Dy_vx = 0, Dx_vz = 0, Dz_vx = 0, Dy_vz = 0, Dz_vy = 0;
    for (int i = 1; i <= 8; i++) {
        double g1 = SHP(1, i, e), g2 = SHP(2, i, e), g3 = SHP(3, i, e);
        int b = 3 * (i - 1);
        Dx_vx = Dx_vx + g1 * va[b];
        Dy_vy = Dy_vy + g2 * va[b + 1];
        Dz_vz = Dz_vz + g3 * va[b + 2];
        Dx_vy = Dx_vy + g1 * va[b + 1];
        Dy_vx = Dy_vx + g2 * va[b];
        Dx_vz = Dx_vz + g1 * va[b + 2];
        Dz_vx = Dz_vx + g3 * va[b];
        Dy_vz = Dy_vz + g2 * va[b + 2];
        Dz_vy = Dz_vy + g3 * va[b + 1];
    }
    const double dt = s->dt;
    double l2m = lam + 2 * miu;
    /* component -> (coefficient, gradient, damping direction), src/assembleGlobalKU.f90:277-311 */
    double coef[15] = {l2m, lam, lam, lam, l2m, lam, lam, lam, l2m, miu, miu, miu, miu, miu, miu};
    double grad[15] = {Dx_vx, Dy_vy, Dz_vz, Dx_vx, Dy_vy, Dz_vz, Dx_vx, Dy_vy, Dz_vz,
                       Dx_vy, Dy_vx, Dx_vz, Dz_vx, Dy_vz, Dz_vy};
    static const int dir[15] = {0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 0, 2, 1, 2};
    for (int c = 0; c < 15; c++) {
        double dd = damps[dir[c]];
        sv[c] = coef[c] * grad[c] + (1 / dt - dd / 2) * sv[c];
        sv[c] = sv[c] / (1 / dt + dd / 2);
    }
    double sxx = sv[0] + sv[1] + sv[2];
    double syy = sv[3] + sv[4] + sv[5];
    double szz = sv[6] + sv[7] + sv[8];
    double sxy = sv[9] + sv[10];
    double sxz = sv[11] + sv[12];
    double syz = sv[13] + sv[14];
    double s0[6];
    for (int k = 0; k < 6; k++) s0[k] = sv[15 + k] + s->rdampk * srt[k];
    double dw = det * s->w;
    for (int i = 1; i <= 8; i++) {
        double g1 = SHP(1, i, e), g2 = SHP(2, i, e), g3 = SHP(3, i, e);
        double *fi = f + (i - 1) * 12;
        fi[0] = fi[0] - dw * g1 * sxx;
        fi[1] = fi[1] - dw * g2 * sxy;
        fi[2] = fi[2] - dw * g3 * sxz;
        fi[3] = fi[3] - dw * g1 * sxy;
        fi[4] = fi[4] - dw * g2 * syy;
        fi[5] = fi[5] - dw * g3 * syz;
        fi[6] = fi[6] - dw * g1 * sxz;
        fi[7] = fi[7] - dw * g2 * syz;
        fi[8] = fi[8] - dw * g3 * szz;
        fi[9] = fi[9] - dw * (g1 * s0[0] + g3 * s0[4] + g2 * s0[5]);
        fi[10] = fi[10] - dw * (g2 * s0[1] + g3 * s0[3] + g1 * s0[5]);
        fi[11] = fi[11] - dw * (g3 * s0[2] + g2 * s0[3] + g1 * s0[4]);
    }
}

/* Force grouping.  0 (default): the reference's, every element's stiffness
 * force added to the nodes first (assembleGlobalKU), then every element's
 * hourglass force mode by mode (hrglss).  1: the grouping of the project's GPU
 * kernels, kept here only to measure what that regrouping alone changes: an
 * element's stiffness and hourglass (C_hg = 1) forces on a node are combined
 * before the node sums its elements -- regular elements subtract mode m's
 * force from their stiffness force mode by mode, PML elements sum the four
 * modes from zero and add them to components 10-12 (or to the collapsed
 * three-component sum, itself formed before the node adds it). */
static int g_grouping = 0;

void eqdo_set_grouping(int32_t g) { g_grouping = g; }

/* hourglass force of mode m (1-based) of element e on its eight nodes,
 * fhr[a][k], as src/hrglss.f90:20-44 forms it */
static void hourglass_mode(const eqdo_scalars *s, const eqdo_state *st, int e, int m, double fhr[8][3])
{
    double dl[3][8];
    for (int i = 1; i <= 8; i++)
        for (int j = 1; j <= 3; j++) dl[j - 1][i - 1] = DISP(j, IEN(i, e)) + s->rdampk * VEL(j, IEN(i, e));
    double phid[3];
    for (int i = 0; i < 3; i++) {
        phid[i] = 0.0;
        for (int j = 1; j <= 8; j++) phid[i] = phid[i] + PHI(j, m, e) * dl[i][j - 1];
    }
    for (int i = 1; i <= 8; i++) {
        double p = PHI(i, m, e);
        fhr[i - 1][0] = p * (SS(1, e) * phid[0] + SS(2, e) * phid[1] + SS(3, e) * phid[2]);
        fhr[i - 1][1] = p * (SS(2, e) * phid[0] + SS(4, e) * phid[1] + SS(5, e) * phid[2]);
        fhr[i - 1][2] = p * (SS(3, e) * phid[0] + SS(5, e) * phid[1] + SS(6, e) * phid[2]);
    }
}

/* ---------------------------------------------------------- assembleGlobalKU
 * src/assembleGlobalKU.f90:3-67, with calcElemMass (src/calcElemMass.f90). */
static void assemble_global_ku(const eqdo_scalars *s, eqdo_state *st)
{
    const int regroup = g_grouping == 1 && s->C_hg == 1;
    const double gterm = (1.0 - (double)s->C_elastic) * s->grav * (s->roumax - (s->gamar + 1.0) * s->rhow) / s->roumax;
    for (int e = 1; e <= s->ne; e++) {
        double al[24], elresf[24], vl[24], dl[24], ex[24];
        for (int a = 1; a <= 8; a++) {
            int n = IEN(a, e);
            for (int k = 1; k <= 3; k++) {
                vl[(a - 1) * 3 + k - 1] = VEL(k, n);
                dl[(a - 1) * 3 + k - 1] = DISP(k, n);
                ex[(a - 1) * 3 + k - 1] = MC(k, n);
                al[(a - 1) * 3 + k - 1] = s->rdampm * VEL(k, n);
            }
            al[(a - 1) * 3 + 2] = al[(a - 1) * 3 + 2] + gterm;
        }
        for (int i = 0; i < 24; i++) elresf[i] = 0.0;
        for (int i = 0; i < 24; i++) elresf[i] = elresf[i] - al[i] * EMASS(i + 1, e);
        int et = st->elemTypeArr[e - 1];
        if (et == 1 || et > 10) {
            double pstrinc;
            double *stress = st->stressArr + st->stressCompIndexArr[e - 1];
            calc_elem_ku(s, st, e, vl, dl, stress, elresf, -st->eledet[e - 1], st->eleporep[e - 1], &pstrinc, ex);
            st->pstrain[e - 1] = st->pstrain[e - 1] + pstrinc;
            if (regroup) {
                for (int m = 1; m <= 4; m++) {
                    double fhr[8][3];
                    hourglass_mode(s, st, e, m, fhr);
                    for (int i = 0; i < 24; i++) elresf[i] = elresf[i] - fhr[i / 3][i % 3];
                }
            }
            for (int i = 1; i <= 8; i++)
                for (int j = 1; j <= 3; j++) {
                    int eq = EQI(ESL(IEN(i, e)) + j);
                    if (eq > 0) FORCE(eq) = FORCE(eq) + elresf[(i - 1) * 3 + j - 1];
                }
        } else if (et == 2) {
            double efPML[96];
            for (int i = 0; i < 96; i++) efPML[i] = 0.0;
            for (int i = 1; i <= 8; i++)
                for (int j = 1; j <= 3; j++) efPML[(i - 1) * 12 + 9 + j - 1] = elresf[(i - 1) * 3 + j - 1];
            double *sv = st->stressArr + st->stressCompIndexArr[e - 1];
            calc_pml_elem_ku(s, st, e, vl, efPML, sv, ex);
            if (regroup) {
                double hg[8][3] = {{0}}, fhr[8][3];
                for (int m = 1; m <= 4; m++) {
                    hourglass_mode(s, st, e, m, fhr);
                    for (int a = 0; a < 8; a++)
                        for (int k = 0; k < 3; k++) hg[a][k] = hg[a][k] - fhr[a][k];
                }
                for (int i = 1; i <= 8; i++) {
                    int n = IEN(i, e);
                    const double *fi = efPML + (i - 1) * 12;
                    if (NDOF(n) == 12) {
                        for (int j = 1; j <= 12; j++) {
                            int eq = EQI(ESL(n) + j);
                            double f = j <= 9 ? fi[j - 1] : fi[j - 1] + hg[i - 1][j - 10];
                            if (eq > 0) FORCE(eq) = FORCE(eq) + f;
                        }
                    } else if (NDOF(n) == 3) {
                        for (int k = 0; k < 3; k++) {
                            int eq = EQI(ESL(n) + k + 1);
                            FORCE(eq) = FORCE(eq) + (fi[3 * k] + fi[3 * k + 1] + fi[3 * k + 2] + fi[9 + k] + hg[i - 1][k]);
                        }
                    }
                }
                continue;
            }
            for (int i = 1; i <= 8; i++) {
                int n = IEN(i, e);
                const double *fi = efPML + (i - 1) * 12;
                if (NDOF(n) == 12) {
                    for (int j = 1; j <= 12; j++) {
                        int eq = EQI(ESL(n) + j);
                        if (eq > 0) FORCE(eq) = FORCE(eq) + fi[j - 1];
                    }
                } else if (NDOF(n) == 3) {
                    int eq = EQI(ESL(n) + 1);
                    FORCE(eq) = FORCE(eq) + fi[0] + fi[1] + fi[2] + fi[9];
                    eq = EQI(ESL(n) + 2);
                    FORCE(eq) = FORCE(eq) + fi[3] + fi[4] + fi[5] + fi[10];
                    eq = EQI(ESL(n) + 3);
                    FORCE(eq) = FORCE(eq) + fi[6] + fi[7] + fi[8] + fi[11];
                }
            }
        }
    }
}

/* ------------------------------------------------------------------- hrglss
 * src/hrglss.f90:3-100 */
static void hourglass(const eqdo_scalars *s, eqdo_state *st)
{
    if (g_grouping == 1 && s->C_hg == 1) return; /* already added per element */
    static const int fi[4][8] = {{1, 1, -1, -1, -1, -1, 1, 1}, {1, -1, -1, 1, -1, 1, 1, -1},
                                 {1, -1, 1, -1, 1, -1, 1, -1}, {1, -1, 1, -1, -1, 1, -1, 1}};
    for (int e = 1; e <= s->ne; e++) {
        double vl[3][8], dl[3][8];
        for (int i = 1; i <= 8; i++)
            for (int j = 1; j <= 3; j++) {
                vl[j - 1][i - 1] = VEL(j, IEN(i, e));
                dl[j - 1][i - 1] = DISP(j, IEN(i, e)) + s->rdampk * vl[j - 1][i - 1];
            }
        if (s->C_hg == 1) {
            for (int m = 1; m <= 4; m++) {
                double phid[3], fhr[8][3];
                for (int i = 0; i < 3; i++) {
                    phid[i] = 0.0;
                    for (int j = 1; j <= 8; j++) phid[i] = phid[i] + PHI(j, m, e) * dl[i][j - 1];
                }
                for (int i = 1; i <= 8; i++) {
                    double p = PHI(i, m, e);
                    fhr[i - 1][0] = p * (SS(1, e) * phid[0] + SS(2, e) * phid[1] + SS(3, e) * phid[2]);
                    fhr[i - 1][1] = p * (SS(2, e) * phid[0] + SS(4, e) * phid[1] + SS(5, e) * phid[2]);
                    fhr[i - 1][2] = p * (SS(3, e) * phid[0] + SS(5, e) * phid[1] + SS(6, e) * phid[2]);
                }
                for (int i = 1; i <= 8; i++) {
                    int n = IEN(i, e);
                    for (int j = 1; j <= 3; j++) {
                        int tag = NDOF(n) == 3 ? ESL(n) + j : ESL(n) + j + 9;
                        int k = EQI(tag);
                        if (k > 0) FORCE(k) = FORCE(k) - fhr[i - 1][j - 1];
                    }
                }
            }
        } else if (s->C_hg == 2) {
            double coef = 0.25 * s->kapa_hg * MAT(e, 3) * MAT(e, 1) * pow(st->eledet[e - 1] * s->w, 2.0 / 3.0);
            double q[3][4], f[24];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 4; j++) {
                    q[i][j] = 0.0;
                    for (int k = 0; k < 8; k++) q[i][j] = q[i][j] + vl[i][k] * (double)fi[j][k];
                }
            for (int i = 0; i < 24; i++) f[i] = 0.0;
            for (int k = 0; k < 8; k++)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 4; j++) f[k * 3 + i] = f[k * 3 + i] - coef * q[i][j] * (double)fi[j][k];
            for (int i = 1; i <= 8; i++) {
                int n = IEN(i, e);
                for (int j = 1; j <= 3; j++) {
                    int tag = NDOF(n) == 3 ? ESL(n) + j : ESL(n) + j + 9;
                    int k = EQI(tag);
                    if (k > 0) FORCE(k) = FORCE(k) + f[(i - 1) * 3 + j - 1];
                }
            }
        }
    }
}

/* --------------------------------------------------------------- friction
 * src/fric.f90:3-95 */
static double slip_weak(double slip, const double *f)
{
    double xmu = 0.0;
    if (fabs(slip) < 1.0e-10)
        xmu = f[0];
    else if (slip < f[2])
        xmu = f[0] - (f[0] - f[1]) * slip / f[2];
    if (slip >= f[2]) xmu = f[1];
    return xmu;
}

static double time_weak(double trupt, const double *f)
{
    if (trupt <= 0.0) return f[0];
    if (trupt < f[4]) return f[0] - (f[0] - f[1]) * trupt / f[4];
    return f[1];
}

void eqdo_ageing_law(double V2, double *theta, const double *f, double dt, double *xmu, double *dxmudv)
{
    double A = f[8], B = f[9], L = f[10], f0 = f[12], V0 = f[11];
    double tmpc = 1.0 / (2.0 * V0) * exp((f0 + B * log(V0 * *theta / L)) / A);
    double tmp = (V2 + 1.0e-30) * tmpc;
    *xmu = A * log(tmp + sqrt(sq(tmp) + 1.0));
    *dxmudv = A * tmpc / sqrt(1.0 + sq(tmp));
    *theta = L / V2 + (*theta - L / V2) * exp(-V2 * dt / L);
}

void eqdo_slip_law(double V2, double *psi, const double *f, double dt, double *xmu, double *dxmudv)
{
    double A = f[8], B = f[9], L = f[10], f0 = f[12], V0 = f[11], fw = f[13], Vw = f[14];
    double tmpc = 1.0 / (2.0 * V0) * exp(*psi / A);
    double tmp = (V2 + 1.0e-30) * tmpc;
    *xmu = A * log(tmp + sqrt(sq(tmp) + 1.0));
    *dxmudv = A * tmpc / sqrt(1.0 + sq(tmp));
    double fLV = f0 - (B - A) * log(V2 / V0);
    /* (V2/Vw)**8: flang expands the integer power as a chain x*(x*(...(x*x))) */
    double r = V2 / Vw, r8 = r * r;
    for (int k = 2; k < 8; k++) r8 = r * r8;
    double fss = fw + (fLV - fw) / pow(1.0 + r8, 0.125);
    double fssa = fss / A;
    double psiss = A * log(2.0 * V0 / V2 * (exp(fssa) - exp(-fssa)) / 2.0);
    *psi = psiss + (*psi - psiss) * exp(-V2 * dt / L);
}

/* ------------------------------------------------------------------ fault
 * src/faulting.f90 */
static void rate_state_normal_stress(double V2, double *theta_pc, double *theta_pc_dot, double tnrm,
                                     const double *f, double dt)
{
    double L = f[10];
    *theta_pc_dot = -(V2 / L * (*theta_pc - fabs(tnrm)));
    *theta_pc = *theta_pc + *theta_pc_dot * dt;
}

static void rsf_nucleation(const eqdo_scalars *s, eqdo_state *st, int i, double T[4], const double R[4])
{
    int sl = NSMP(1, i);
    double dtau = 0.0;
    double radius = sqrt(sq(MC(1, sl) - s->xsource) + sq(MC(2, sl) - s->ysource) + sq(MC(3, sl) - s->zsource));
    double F = 0.0, G = 1.0, t = st->timeElapsed;
    if (radius < s->nucR) F = exp(sq(radius) / (sq(radius) - sq(s->nucR)));
    if (t <= s->nucT) G = exp(sq(t - s->nucT) / (t * (t - 2.0 * s->nucT)));
    if (s->TPV == 105 || s->TPV == 104) {
        dtau = s->nucdtau0 * F * G;
    } else if (s->TPV == 2802) {
        if (st->nt == 1) {
            FRIC(81, i) = s->nucdtau0;
            double ttao = sqrt(sq(T[1]) + sq(T[2]));
            double back = sqrt(sq(R[1] + FRIC(26, i)) + sq(R[2] + FRIC(27, i)));
            FRIC(20, i) = FRIC(9, i) * log(2.0 * FRIC(12, i) / back * sinh(ttao / fabs(T[0]) / FRIC(9, i)));
            FRIC(23, i) = fabs(T[0]);
        }
        dtau = FRIC(81, i) * F * G;
    }
    T[1] = T[1] + dtau;
}

static void sw_tw_nucleation(const eqdo_scalars *s, eqdo_state *st, int i, double *fricCoeff)
{
    int sl = NSMP(1, i);
    double radius = sqrt(sq(MC(1, sl) - s->xsource) + sq(MC(2, sl) - s->ysource) + sq(MC(3, sl) - s->zsource));
    double tr = 1.0e9, tc, t = st->timeElapsed;
    if (radius <= s->nucR) {
        if (s->TPV == 201 || s->TPV == 36 || s->TPV == 37)
            tr = (radius + 0.081 * s->nucR * (1.0 / (1.0 - sq(radius / s->nucR)) - 1.0)) / (0.7 * 3464.0);
        if (s->TPV == 202) tr = radius / s->nucRuptVel;
    }
    tc = 1.0;
    if (t < tr)
        tc = 0.0;
    else if ((t < (tr + FRIC(5, i))) && (t >= tr))
        tc = (t - tr) / FRIC(5, i);
    double cand = FRIC(1, i) + (FRIC(2, i) - FRIC(1, i)) * tc;
    *fricCoeff = cand < *fricCoeff ? cand : *fricCoeff;
}

static void newton_raphson(const eqdo_scalars *s, eqdo_state *st, int i, double *v_trial, double *taoc_new,
                           double *state0, double *thetaPc0, double *thetaPcDot, const double T[4],
                           const double trial[4], double Tc)
{
    double stateTmp = 0.0, thetaPcTmp = 0.0, xmu, dxmudv, rsfeq, drsfeqdv;
    const double *f = &FRIC(1, i);
    for (int iv = 1; iv <= 20; iv++) {
        stateTmp = *state0;
        if (s->friclaw == 3)
            eqdo_ageing_law(*v_trial, &stateTmp, f, s->dt, &xmu, &dxmudv);
        else
            eqdo_slip_law(*v_trial, &stateTmp, f, s->dt, &xmu, &dxmudv);
        if (s->friclaw < 5) {
            thetaPcTmp = *thetaPc0;
            rate_state_normal_stress(*v_trial, &thetaPcTmp, thetaPcDot, T[0], f, s->dt);
            *taoc_new = xmu * thetaPcTmp;
            rsfeq = *v_trial + Tc * (*taoc_new * 0.5 - trial[3]);
            drsfeqdv = 1.0 + Tc * (dxmudv * thetaPcTmp) * 0.5;
        } else {
            double tn = T[0] < 0.0 ? T[0] : 0.0;
            *taoc_new = FRIC(4, i) - xmu * tn;
            rsfeq = *v_trial + Tc * (*taoc_new * 0.5 - trial[3]);
            drsfeqdv = 1.0 + Tc * (-dxmudv * tn) * 0.5;
        }
        if (fabs(rsfeq / drsfeqdv) < 1.0e-14 * fabs(*v_trial) && fabs(rsfeq) < 1.0e-6 * fabs(*v_trial)) break;
        double nv = *v_trial - rsfeq / drsfeqdv;
        if (nv <= 0.0)
            *v_trial = *v_trial / 2.0;
        else
            *v_trial = nv;
    }
    if (s->TPV == 105 && *v_trial < FRIC(46, i)) *v_trial = FRIC(46, i);
    *state0 = stateTmp;
    *thetaPc0 = thetaPcTmp;
}

static void faulting(const eqdo_scalars *s, eqdo_state *st)
{
    const double dt = s->dt;
    const double Ce = (double)s->C_elastic;
    for (int i = 1; i <= s->nf; i++) {
        double slip[4], rate[4], T[4], T0[3], q[3][2][3], nsd[3][2][3];
        int nodes[2] = {NSMP(1, i), NSMP(2, i)};
        /* getNsdSlipSliprateTraction, src/faulting.f90:54-134 (dtau is always 0) */
        T0[0] = FRIC(7, i);
        T0[1] = FRIC(8, i) + 0.0;
        T0[2] = FRIC(49, i);
        double ms = st->fnms[nodes[0] - 1], mm = st->fnms[nodes[1] - 1];
        double totalMass = (ms + mm) * st->arn[i - 1];
        for (int j = 0; j < 2; j++)
            for (int k = 1; k <= 3; k++) {
                q[k - 1][j][0] = FORCE(EQI(ESL(nodes[j]) + k));
                q[k - 1][j][1] = VEL(k, nodes[j]);
                q[k - 1][j][2] = DISP(k, nodes[j]);
            }
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 2; k++) {
                nsd[0][k][j] = q[0][k][j] * UN(1, i) + q[1][k][j] * UN(2, i) + q[2][k][j] * UN(3, i);
                nsd[1][k][j] = q[0][k][j] * US(1, i) + q[1][k][j] * US(2, i) + q[2][k][j] * US(3, i);
                nsd[2][k][j] = q[0][k][j] * UD(1, i) + q[1][k][j] * UD(2, i) + q[2][k][j] * UD(3, i);
            }
        for (int j = 0; j < 3; j++) slip[j] = nsd[j][1][2] - nsd[j][0][2];
        slip[3] = sqrt(sq(slip[0]) + sq(slip[1]) + sq(slip[2]));
        for (int j = 0; j < 3; j++) rate[j] = nsd[j][1][1] - nsd[j][0][1];
        rate[3] = sqrt(sq(rate[0]) + sq(rate[1]) + sq(rate[2]));
        FRIC(71, i) = slip[1];
        FRIC(72, i) = slip[2];
        FRIC(73, i) = slip[0];
        FRIC(74, i) = rate[1];
        FRIC(75, i) = rate[2];
        if (rate[3] > FRIC(76, i)) FRIC(76, i) = rate[3];
        FRIC(77, i) = FRIC(77, i) + rate[3] * dt;
        T[0] = (ms * mm * ((nsd[0][1][1] - nsd[0][0][1]) + (nsd[0][1][2] - nsd[0][0][2]) / dt) / dt +
                ms * nsd[0][1][0] - mm * nsd[0][0][0]) / totalMass + T0[0] * Ce;
        T[1] = (ms * mm * (nsd[1][1][1] - nsd[1][0][1]) / dt + ms * nsd[1][1][0] - mm * nsd[1][0][0]) / totalMass +
               T0[1] * Ce;
        T[2] = (ms * mm * (nsd[2][1][1] - nsd[2][0][1]) / dt + ms * nsd[2][1][0] - mm * nsd[2][0][0]) / totalMass +
               T0[2] * Ce;
        if (s->friclaw >= 3 && s->C_nuclea == 1 && 1 == s->nucfault) rsf_nucleation(s, st, i, T, rate);
        T[3] = sqrt(sq(T[1]) + sq(T[2]));

        if (s->friclaw <= 2) {
            /* solveSWTW, src/faulting.f90:136-188 */
            double mu = 0.0, xyzT[3], xyzT0[3];
            if (s->friclaw == 1)
                mu = slip_weak(FRIC(77, i), &FRIC(1, i));
            else if (s->friclaw == 2)
                mu = time_weak(st->timeElapsed - st->fnft[i - 1], &FRIC(1, i));
            if (s->C_nuclea == 1 && 1 == s->nucfault) sw_tw_nucleation(s, st, i, &mu);
            double eff = (T[0] + FRIC(6, i)) > 0 ? 0.0 : T[0] + FRIC(6, i);
            double trialShear = FRIC(4, i) - mu * eff;
            if (T[3] > trialShear) {
                T[1] = T[1] * trialShear / T[3];
                T[2] = T[2] * trialShear / T[3];
            }
            for (int j = 1; j <= 3; j++) {
                xyzT[j - 1] = (T[0] * UN(j, i) + T[1] * US(j, i) + T[2] * UD(j, i)) * st->arn[i - 1];
                xyzT0[j - 1] = (T0[0] * UN(j, i) + T0[1] * US(j, i) + T0[2] * UD(j, i)) * st->arn[i - 1];
            }
            for (int j = 1; j <= 3; j++) {
                int eqs = EQI(ESL(nodes[0]) + j), eqm = EQI(ESL(nodes[1]) + j);
                FORCE(eqs) = FORCE(eqs) + xyzT[j - 1] - xyzT0[j - 1] * Ce;
                FORCE(eqm) = FORCE(eqm) - xyzT[j - 1] + xyzT0[j - 1] * Ce;
            }
            for (int j = 0; j < 3; j++) FRIC(78 + j, i) = T[j];
        } else {
            /* solveRSF, src/faulting.f90:190-327 */
            double trial[4], acc[3], xyzAcc[3], xyzR[3];
            if (s->friclaw == 5)
                T[0] = T[0] + FRIC(51, i);
            else
                T[0] = T[0] + FRIC(6, i);
            if (s->insertFaultType > 0 && s->C_elastic == 1) {
                const double max_norm = -40.0e6, min_norm = -10.0e6;
                if (T[0] >= min_norm)
                    T[0] = min_norm;
                else if (T[0] <= max_norm)
                    T[0] = max_norm;
            }
            if (T[0] > 0.0) T[0] = 0.0;
            for (int j = 0; j < 3; j++) {
                slip[j] = slip[j] + FRIC(25 + j, i) * st->timeElapsed;
                rate[j] = rate[j] + FRIC(25 + j, i);
            }
            slip[3] = sqrt(sq(slip[1]) + sq(slip[2]));
            rate[3] = sqrt(sq(rate[1]) + sq(rate[2]));
            double v_trial = rate[3];
            double theta_pc_tmp = FRIC(23, i), theta_pc_dot, xmu, dxmudv;
            rate_state_normal_stress(v_trial, &FRIC(23, i), &theta_pc_dot, T[0], &FRIC(1, i), dt);
            FRIC(24, i) = theta_pc_dot;
            double statetmp = FRIC(20, i);
            if (s->friclaw == 3)
                eqdo_ageing_law(v_trial, &FRIC(20, i), &FRIC(1, i), dt, &xmu, &dxmudv);
            else
                eqdo_slip_law(v_trial, &FRIC(20, i), &FRIC(1, i), dt, &xmu, &dxmudv);
            double taoc_old = s->friclaw == 5 ? FRIC(4, i) - xmu * T[0] : xmu * theta_pc_tmp;
            double mr = mm * ms / (mm + ms);
            double Tc = st->arn[i - 1] * dt / mr;
            for (int j = 1; j <= 2; j++)
                trial[j] = T[j] - taoc_old * 0.5 * (rate[j] / rate[3]) + FRIC(25 + j, i) / Tc;
            trial[3] = sqrt(sq(trial[1]) + sq(trial[2]));
            double taoc_new = 0.0;
            newton_raphson(s, st, i, &v_trial, &taoc_new, &statetmp, &theta_pc_tmp, &theta_pc_dot, T, trial, Tc);
            FRIC(20, i) = statetmp;
            FRIC(23, i) = theta_pc_tmp;
            for (int j = 1; j <= 2; j++)
                T[j] = taoc_old * 0.5 * (rate[j] / rate[3]) + taoc_new * 0.5 * (trial[j] / trial[3]);
            for (int j = 0; j < 3; j++) FRIC(78 + j, i) = T[j];
            FRIC(47, i) = v_trial;
            FRIC(48, i) = sqrt(sq(T[1]) + sq(T[2]));
            if (st->tpHist) {  /* src/faulting.f90:301-302 */
                st->tpHist[((size_t)(st->nt - 1) * s->nf + i - 1) * 2 + 0] = FRIC(47, i);
                st->tpHist[((size_t)(st->nt - 1) * s->nf + i - 1) * 2 + 1] = FRIC(48, i);
            }
            acc[0] = -rate[0] / dt - slip[0] / dt / dt;
            acc[1] = (v_trial * (trial[1] / trial[3]) - rate[1]) / dt;
            acc[2] = (v_trial * (trial[2] / trial[3]) - rate[2]) / dt;
            for (int j = 1; j <= 3; j++) {
                xyzAcc[j - 1] = acc[0] * UN(j, i) + acc[1] * US(j, i) + acc[2] * UD(j, i);
                xyzR[j - 1] = FORCE(EQI(ESL(nodes[0]) + j)) + FORCE(EQI(ESL(nodes[1]) + j));
            }
            for (int j = 1; j <= 3; j++) {
                FORCE(EQI(ESL(nodes[0]) + j)) = (-xyzAcc[j - 1] + xyzR[j - 1] / mm) * mr;
                FORCE(EQI(ESL(nodes[1]) + j)) = (xyzAcc[j - 1] + xyzR[j - 1] / ms) * mr;
                FRIC(31 + j - 1, i) = VEL(j, nodes[1]) + (xyzAcc[j - 1] + xyzR[j - 1] / ms) * dt;
                FRIC(34 + j - 1, i) = VEL(j, nodes[0]) + (-xyzAcc[j - 1] + xyzR[j - 1] / mm) * dt;
            }
        }
        /* storeOnFaultStationQuantSCEC, src/faulting.f90:518-541 */
        for (int j = 1; j <= s->non; j++) {
            if (st->anonfs[(j - 1) * 3 + 0] == i && st->anonfs[(j - 1) * 3 + 2] == 1) {
                double *h = st->onFault + ((size_t)(j - 1) * s->nstep + (st->nt - 1)) * 12;
                h[0] = st->timeElapsed;
                h[1] = rate[1];
                h[2] = rate[2];
                h[3] = FRIC(20, i);
                h[4] = slip[1];
                h[5] = slip[2];
                h[6] = slip[0];
                h[7] = T[1];
                h[8] = T[2];
                h[9] = T[0];
                h[10] = FRIC(51, i) + FRIC(42, i);
                h[11] = FRIC(52, i);
            }
        }
        /* storeRuptureTime, src/faulting.f90:329-341 */
        if (st->fnft[i - 1] > 5000.0 && rate[3] >= s->slipRateThres) st->fnft[i - 1] = st->timeElapsed;
    }
}

/* thermop, src/thermop.f90:1-40: pore pressure change and temperature of
 * every pair from the history of slip rate and shear traction up to the
 * previous step (onFaultTPHist(1:2, i, 1:nt-1)). */
#define TPH(k, i, j) st->tpHist[((size_t)((j) - 1) * s->nf + (i) - 1) * 2 + (k) - 1]
static void thermop(const eqdo_scalars *s, eqdo_state *st)
{
    const double dt = s->dt, h = s->fric_tp_h, nt = st->nt;
    for (int i = 1; i <= s->nf; i++) {
        const double gama = FRIC(19, i) / FRIC(18, i);
        const double omega = FRIC(16, i), kapa = FRIC(17, i);
        double tmp = 0.0;
        for (int j = 1; j <= nt - 1; j++) {
            double ker = -kapa / (omega - kapa) / sqrt(4.0 * kapa * (double)(nt - j) * dt + 2.0 * (h * h));
            ker = ker + omega / (omega - kapa) / sqrt(4.0 * omega * (double)(nt - j) * dt + 2.0 * (h * h));
            tmp = tmp + fabs(TPH(2, i, j)) * TPH(1, i, j) * ker * dt;
        }
        const double patnode = tmp * gama / sqrt(s->pi);
        tmp = 0.0;
        for (int j = 1; j <= nt - 1; j++) {
            const double ker = 1.0 / sqrt(4.0 * kapa * (double)(nt - j) * dt + 2.0 * (h * h));
            tmp = tmp + fabs(TPH(2, i, j)) * TPH(1, i, j) * ker * dt;
        }
        const double tatnode = tmp / FRIC(18, i) / sqrt(s->pi);
        FRIC(51, i) = patnode;
        FRIC(52, i) = tatnode + FRIC(41, i);
    }
}

/* ------------------------------------------------------------------- driver
 * src/driver.f90:3-36 (one rank: MPI4NodalQuant is the identity). */
int eqdo_run(const eqdo_scalars *s, eqdo_state *st, int32_t nsteps)
{
    for (int k = 0; k < nsteps; k++) {
        st->nt = st->nt + 1;
        st->timeElapsed = st->timeElapsed + s->dt;
        if (vel_disp_update(s, st)) return 1;
        store_off_fault(s, st);
        memset(st->nodalForceArr, 0, sizeof(double) * (size_t)s->neq);
        assemble_global_ku(s, st);
        hourglass(s, st);
        if (s->friclaw == 5) thermop(s, st);
        faulting(s, st);
        for (int i = 0; i < s->neq; i++) st->nodalForceArr[i] = st->nodalForceArr[i] / st->nodalMassArr[i];
    }
    return 0;
}
