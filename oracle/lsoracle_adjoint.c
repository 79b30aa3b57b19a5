/*
 * lsoracle_adjoint.c -- the differentiable form of the update, restated on the
 * host (TEST INFRASTRUCTURE ONLY, see lsoracle.h).
 *
 * No reference counterpart: the reference only runs the forward rule.  Here
 * the forward rule of UpdateLevelSetVolume.cu:222-342 (as restated in
 * lsoracle_update.inc) is taken with its branch decisions held fixed (the two
 * clamps, the connectivity gate of :278-292, a vanishing |grad phi|, the side
 * of |I - T| in ComputeDensityTerm :335-342), which
 * makes the map from (phi0, T, eps, alpha) to the final phi smooth, and its
 * adjoint is written out in double precision.
 *
 * One iteration maps the buffer pair (A, B) (read, write; Segmenter.cpp:
 * 1339-1351 swaps them) to A'(v) = f_v(A) for v in V_k, A'(v) = B(v)
 * otherwise, B' = A.  Its adjoint, with (Abar', Bbar') the adjoints of the
 * outputs, is Abar = Bbar' + sum_{v in V_k} J_v^T Abar'(v) and Bbar = Abar'
 * outside V_k (0 on V_k).
 *
 * All arrays are padded and tiled (tau order), like the rest of the oracle.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lsoracle.h"

/* decision bits, shared with the device code's recording; DEC_BELOW (float
   sources) is the side of the window's |I - T| */
enum { DEC_LO1 = 1, DEC_HI1 = 2, DEC_LO2 = 4, DEC_HI2 = 8, DEC_CONN = 16, DEC_FLAT = 32, DEC_BELOW = 64 };

static int32_t source_int(int32_t dtype, const void* src, uint32_t t)
{
    switch (dtype) {
    case LSO_U8: return ((const uint8_t*)src)[t];
    case LSO_I8: return ((const int8_t*)src)[t];
    case LSO_U16: return ((const uint16_t*)src)[t];
    case LSO_I16: return ((const int16_t*)src)[t];
    case LSO_U32: return (int32_t)((const uint32_t*)src)[t];
    default: return ((const int32_t*)src)[t];
    }
}

/* The decisions of the FP32 update of every voxel of V (the same comparisons
   as lsoracle_update.inc's float instantiation), one byte per entry. */
void lso_decisions_f32(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src, const float* R,
                       float T, float eps, float alpha, float dt, const uint32_t* V, int64_t n, uint8_t* dec)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        const uint32_t t = V[p];
        const float unz = R[lso_neighbor(X, Y, Z, t, LSO_NZ)], u1 = R[lso_neighbor(X, Y, Z, t, LSO_NY)];
        const float u3 = R[lso_neighbor(X, Y, Z, t, LSO_NX)], u4 = R[t];
        const float u5 = R[lso_neighbor(X, Y, Z, t, LSO_PX)], u7 = R[lso_neighbor(X, Y, Z, t, LSO_PY)];
        const float upz = R[lso_neighbor(X, Y, Z, t, LSO_PZ)];
        const float dcx = (u5 - u3) / 2.0f, dcy = (u7 - u1) / 2.0f, dcz = (upz - unz) / 2.0f;
        const float sdx = (u5 - u4) - (u4 - u3), sdy = (u7 - u4) - (u4 - u1), sdz = (upz - u4) - (u4 - unz);
        const float tol = 0.1f;
        const int gate = alpha > 0.0f && fabsf(upz - unz) <= tol && fabsf(unz - u1) <= tol &&
                         fabsf(u1 - u3) <= tol && fabsf(u3 - u5) <= tol && fabsf(u5 - u7) <= tol &&
                         fabsf(u7 - upz) <= tol;
        const float conn = gate ? (upz + u1 + u3 + u5 + u7 + unz) / 6.0f - u4 : 0.0f;
        float dens;
        if (dtype == LSO_F32) {
            const float d = ((const float*)src)[t] - T;
            dens = -((-1.0f / eps) * (d < 0.0f ? -d : d) + 1.0f);
        } else {
            const int32_t d = source_int(dtype, src, t) - (int32_t)T;
            dens = -((-1.0f / (float)(uint32_t)(int32_t)eps) * (float)(d < 0 ? -d : d) + 1.0f);
        }
        float dta = dt;
        if (dt < 0.0f) {
            dta = -dt;
            dens = -dens;
        }
        const float speed = ((1.0f - alpha) * dens) + (alpha * (sdx + sdy + sdz));
        const float g = sqrtf(dcx * dcx + dcy * dcy + dcz * dcz);
        const float x1 = u4 + ((dta * speed * g) + conn);
        const float c1 = x1 > -1.0f ? (x1 < 1.0f ? x1 : 1.0f) : -1.0f;
        const float x2 = ((1.0f + 0.01f) * c1) + 0.01f;
        uint8_t d8 = 0;
        if (!(x1 > -1.0f)) d8 |= DEC_LO1;
        else if (!(x1 < 1.0f)) d8 |= DEC_HI1;
        if (!(x2 > -1.0f)) d8 |= DEC_LO2;
        else if (!(x2 < 1.0f)) d8 |= DEC_HI2;
        if (gate) d8 |= DEC_CONN;
        if (g == 0.0f) d8 |= DEC_FLAT;
        if (dtype == LSO_F32 && ((const float*)src)[t] - T < 0.0f) d8 |= DEC_BELOW;
        dec[p] = d8;
    }
}

/* intermediates of one replayed update in double */
typedef struct {
    uint32_t nb[7];
    double u[7];  /* taps in list order 0:-z 1:-y 2:-x 3:self 4:+x 5:+y 6:+z */
    double dcx, dcy, dcz, curv, g, dens, speed, absdi, sgn, x2;
} upd64;

static void eval64(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src, const double* R, uint32_t t,
                   double T, double eps, double alpha, double dt, uint8_t dec, upd64* o)
{
    for (int e = 0; e < 7; ++e) {
        o->nb[e] = e == LSO_SELF ? t : lso_neighbor(X, Y, Z, t, e);
        o->u[e] = R[o->nb[e]];
    }
    const double unz = o->u[0], u1 = o->u[1], u3 = o->u[2], u4 = o->u[3], u5 = o->u[4], u7 = o->u[5],
                 upz = o->u[6];
    o->dcx = (u5 - u3) / 2.0;
    o->dcy = (u7 - u1) / 2.0;
    o->dcz = (upz - unz) / 2.0;
    o->curv = ((u5 - u4) - (u4 - u3)) + ((u7 - u4) - (u4 - u1)) + ((upz - u4) - (u4 - unz));
    double D;
    if (dtype == LSO_F32) {
        /* |I - T| on its recorded side */
        const double di = (double)((const float*)src)[t] - T;
        o->sgn = (dec & DEC_BELOW) ? -1.0 : 1.0;
        o->absdi = o->sgn * di;
        D = 1.0 - o->absdi / eps;
    } else {
        const int32_t di = source_int(dtype, src, t) - (int32_t)T;
        o->absdi = fabs((double)di);
        o->sgn = 0.0;
        D = 1.0 - o->absdi / (double)(uint32_t)(int32_t)eps;
    }
    o->dens = dt < 0 ? D : -D;
    o->speed = (1.0 - alpha) * o->dens + alpha * o->curv;
    o->g = sqrt(o->dcx * o->dcx + o->dcy * o->dcy + o->dcz * o->dcz);
    const double conn = (dec & DEC_CONN) ? (upz + u1 + u3 + u5 + u7 + unz) / 6.0 - u4 : 0.0;
    const double x1 = u4 + fabs(dt) * o->speed * o->g + conn;
    const double c1 = (dec & DEC_LO1) ? -1.0 : ((dec & DEC_HI1) ? 1.0 : x1);
    const double x2 = (double)(1.0f + 0.01f) * c1 + (double)0.01f;
    o->x2 = (dec & DEC_LO2) ? -1.0 : ((dec & DEC_HI2) ? 1.0 : x2);
}

/* Replay in double: both buffers := phi0, `iters` updates over the recorded
   lists with the recorded decisions.  phi_out gets the final read buffer;
   A_hist (iters x N, may be NULL) the read buffer entering each iteration. */
void lso_replay_f64(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src, double T, double eps,
                    double alpha, double dt, int32_t iters, const uint32_t* n, const uint32_t* lists,
                    const uint8_t* dec, const double* phi0, double* phi_out, double* A_hist)
{
    const int64_t N = (int64_t)X * Y * Z;
    double* A = (double*)malloc(N * sizeof(double));
    double* B = (double*)malloc(N * sizeof(double));
    memcpy(A, phi0, N * sizeof(double));
    memcpy(B, phi0, N * sizeof(double));
    int64_t off = 0;
    for (int32_t k = 0; k < iters; ++k) {
        if (A_hist) memcpy(A_hist + (int64_t)k * N, A, N * sizeof(double));
        const int64_t nk = n[k];
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < nk; ++p) {
            upd64 o;
            eval64(X, Y, Z, dtype, src, A, lists[off + p], T, eps, alpha, dt, dec[off + p], &o);
            B[lists[off + p]] = o.x2;
        }
        off += nk;
        double* tmp = A;
        A = B;
        B = tmp;
    }
    memcpy(phi_out, A, N * sizeof(double));
    free(A);
    free(B);
}

/* Adjoint in double of lso_replay_f64 with A_hist from it: dphiK = dL/d(final
   read buffer) -> dphi0 = dL/dphi0 (both buffers started at phi0), dtheta =
   dL/dT, dL/deps, dL/dalpha (T and eps only for float sources). */
void lso_adjoint_f64(int32_t X, int32_t Y, int32_t Z, int32_t dtype, const void* src, double T, double eps,
                     double alpha, double dt, int32_t iters, const uint32_t* n, const uint32_t* lists,
                     const uint8_t* dec, const double* A_hist, const double* dphiK, double* dphi0, double* dtheta)
{
    const int64_t N = (int64_t)X * Y * Z;
    double* Ab = (double*)malloc(N * sizeof(double));
    double* Bb = (double*)calloc(N, sizeof(double));
    memcpy(Ab, dphiK, N * sizeof(double));
    int64_t* offs = (int64_t*)malloc((iters + 1) * sizeof(int64_t));
    offs[0] = 0;
    for (int32_t k = 0; k < iters; ++k) offs[k + 1] = offs[k] + n[k];
    double th[3] = {0, 0, 0};
    const double rescale = (double)(1.0f + 0.01f);
    for (int32_t k = iters - 1; k >= 0; --k) {
        const double* A = A_hist + (int64_t)k * N;
        /* sequential: the scatter into Bb must not race */
        for (int64_t p = 0; p < (int64_t)n[k]; ++p) {
            const uint32_t t = lists[offs[k] + p];
            const uint8_t d = dec[offs[k] + p];
            const double gbar = Ab[t];
            Ab[t] = 0.0;
            if ((d & (DEC_LO1 | DEC_HI1 | DEC_LO2 | DEC_HI2)) || gbar == 0.0) continue;
            upd64 o;
            eval64(X, Y, Z, dtype, src, A, t, T, eps, alpha, dt, d, &o);
            const double s = gbar * rescale;
            const double sp = s * fabs(dt) * o.g;     /* dL/dspeed */
            const double gs = s * fabs(dt) * o.speed; /* dL/d|grad phi| */
            const double gc = sp * alpha;             /* dL/dcurvature */
            const double cn = (d & DEC_CONN) ? s / 6.0 : 0.0;
            double gx = 0, gy = 0, gz = 0;
            if (!(d & DEC_FLAT) && o.g > 0) {
                gx = 0.5 * gs * o.dcx / o.g;
                gy = 0.5 * gs * o.dcy / o.g;
                gz = 0.5 * gs * o.dcz / o.g;
            }
            Bb[o.nb[0]] += gc + cn - gz;
            Bb[o.nb[1]] += gc + cn - gy;
            Bb[o.nb[2]] += gc + cn - gx;
            Bb[t] += s - 6.0 * gc - ((d & DEC_CONN) ? s : 0.0);
            Bb[o.nb[4]] += gc + cn + gx;
            Bb[o.nb[5]] += gc + cn + gy;
            Bb[o.nb[6]] += gc + cn + gz;
            th[2] += sp * (o.curv - o.dens);
            if (dtype == LSO_F32) {
                const double dd = dt < 0 ? 1.0 : -1.0; /* dens = dd * D */
                th[0] += sp * (1.0 - alpha) * dd * (o.sgn / eps);
                th[1] += sp * (1.0 - alpha) * dd * (o.absdi / (eps * eps));
            }
        }
        double* tmp = Ab; /* Abar_k is the scattered buffer, Bbar_k the cleared one */
        Ab = Bb;
        Bb = tmp;
    }
    for (int64_t i = 0; i < N; ++i) dphi0[i] = Ab[i] + Bb[i];
    memcpy(dtheta, th, sizeof(th));
    free(Ab);
    free(Bb);
    free(offs);
}
