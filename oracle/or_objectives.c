/*
 * or_objectives.c -- the eight continuous objectives of the oracle.
 * TEST INFRASTRUCTURE ONLY (see emtf_oracle.h).
 *
 * Follows the reference evaluator `evaluation`, proposed model/SG-EMTF_DE.cu:834-1069
 * (GA copy SG-EMTF.cu:898-1111), with its transform z_d = M_d (x_d + b_d)
 * (SG-EMTF_DE.cu:856-1054, written here as M (x - o) with o = -b) and the
 * rotated transform z = Q (x - o) of the benchmark configurations.
 * Function ids 0..6 are the reference's case labels (Sphere, Rosenbrock,
 * Ackley, Rastrigin, Griewank, Weierstrass, Schwefel); 7 is the CEC
 * high-conditioned Elliptic, which the reference does not have.
 *
 * FP32 forms are the cancellation-free restatements of the same functions:
 *   Rastrigin   z^2 - 10 cos(2 pi z) + 10   = z^2 + 20 sin^2(pi z)
 *   Ackley      20 + e - 20 e^{-0.2 r} - e^{C/D} = -20 expm1(-0.2 r) - e expm1(-2 S/D),
 *               r = sqrt(sum z^2 / D), S = sum sin^2(pi z)
 *   Griewank    1 - prod cos(z_j / sqrt(j+1)) = q_D with q_{k+1} = q_k + a_k (1 - q_k),
 *               a_k = 2 sin^2(z_k / (2 sqrt(k+1)))
 *   Weierstrass sum_k 0.5^k [cos(2 pi 3^k (z + 1/2)) - cos(pi 3^k)] = sum_k 2^{1-k} sin^2(pi 3^k z)
 * so that each equals zero at its optimum.  The FP64 evaluator below uses
 * the textbook forms, which makes it an independent check of the rewrite.
 */
#include "emtf_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SCHWEFEL_C 418.9828872724338
#define SCHWEFEL_CF 418.982887f
#define SCHWEFEL_REF_CF 418.9829f   /* the reference's constant, SG-EMTF_DE.cu:1061 */
#define E_F 2.71828175f
#define E_REF 2.718282               /* the reference's E (a double literal), SG-EMTF_DE.cu:48 */

void or_func_table(int32_t func, int32_t dim, float* table) {
    const float w_ref = (func == OR_FUNC_REF_WEIERSTRASS) ? or_weier_ref_w() : 0.0f;
    for (int j = 0; j < dim; ++j) {
        double v = 0.0;
        if (func == OR_FUNC_ELLIPTIC) v = (dim > 1) ? pow(1.0e6, (double)j / (double)(dim - 1)) : 1.0;
        else if (func == OR_FUNC_GRIEWANK) v = 1.0 / (2.0 * M_PI * sqrt((double)(j + 1)));
        table[j] = (float)v;
        if (func == OR_FUNC_SCHWEFEL) table[j] = SCHWEFEL_CF;
        else if (func == OR_FUNC_REF_SCHWEFEL) table[j] = SCHWEFEL_REF_CF;
        else if (func == OR_FUNC_REF_GRIEWANK) table[j] = sqrtf((float)(j + 1));
        else if (func == OR_FUNC_REF_GRIEWANK_GA) table[j] = INFINITY;
        else if (func == OR_FUNC_REF_WEIERSTRASS) table[j] = w_ref;
    }
}

/* z from x under the task transform (contract arithmetic). */
static void transform32(const or_task* t, const float* x, float* w, float* z) {
    const int D = t->dim;
    for (int d = 0; d < D; ++d) w[d] = x[d] - t->shift[d];
    if (t->xform == OR_XFORM_SHIFT) {
        for (int j = 0; j < D; ++j) z[j] = w[j];
    } else if (t->xform == OR_XFORM_DIAG) {
        for (int j = 0; j < D; ++j) z[j] = t->diag[j] * w[j];
    } else {
        const int Dp = (D + 1) & ~1;
        for (int j = 0; j < D; ++j) {
            float acc = 0.0f;
            const float* q = t->rot + (size_t)j * D;
            for (int d = 0; d < D; ++d) acc = fmaf(q[d], w[d], acc);
            for (int d = D; d < Dp; ++d) acc = fmaf(0.0f, 0.0f, acc);
            z[j] = acc;
        }
    }
}

#define ACC_A(j) ((((j) >> 2) & 1) == 0)

float or_finish32(const or_task* t, const float* z, const float* tab) {
    const int D = t->dim;
    float A = 0.0f, B = 0.0f, A2 = 0.0f, B2 = 0.0f;
    switch (t->func) {
    case OR_FUNC_SPHERE:
        for (int j = 0; j < D; ++j) { float v = z[j] * z[j]; if (ACC_A(j)) A = A + v; else B = B + v; }
        return A + B;
    case OR_FUNC_ROSENBROCK:
        for (int j = 0; j + 1 < D; ++j) {
            float zz = z[j] * z[j];
            float a = zz - z[j + 1];
            float b = z[j] - 1.0f;
            float t1 = 100.0f * a;
            t1 = t1 * a;
            float b2 = b * b;
            float v = t1 + b2;
            if (ACC_A(j)) A = A + v; else B = B + v;
        }
        return A + B;
    case OR_FUNC_ACKLEY: {
        for (int j = 0; j < D; ++j) {
            float v = z[j] * z[j];
            float s = or_sinpi_sq(z[j]);
            if (ACC_A(j)) { A = A + v; A2 = A2 + s; } else { B = B + v; B2 = B2 + s; }
        }
        float Df = (float)D;
        float S = A + B, C = A2 + B2;
        float q = S / Df;
        float r = sqrtf(q);
        float a = -0.2f * r;
        float e1 = or_expm1(a);
        float t1 = -20.0f * e1;
        float c2 = C / Df;
        float b = -2.0f * c2;
        float e2 = or_expm1(b);
        float t2 = E_F * e2;
        return t1 - t2;
    }
    case OR_FUNC_RASTRIGIN:
        for (int j = 0; j < D; ++j) {
            float zz = z[j] * z[j];
            float s = or_sinpi_sq(z[j]);
            float s20 = 20.0f * s;
            float v = zz + s20;
            if (ACC_A(j)) A = A + v; else B = B + v;
        }
        return A + B;
    case OR_FUNC_GRIEWANK: {
        float qA = 0.0f, qB = 0.0f;
        for (int j = 0; j < D; ++j) {
            float v = z[j] * z[j];
            float tt = z[j] * tab[j];
            float s = or_sinpi_sq(tt);
            float a = 2.0f * s;
            if (ACC_A(j)) { A = A + v; float om = 1.0f - qA; float aq = a * om; qA = qA + aq; }
            else { B = B + v; float om = 1.0f - qB; float aq = a * om; qB = qB + aq; }
        }
        float om = 1.0f - qA;
        float tq = qB * om;
        float Q = qA + tq;
        float S = A + B;
        float s4 = S / 4000.0f;
        return s4 + Q;
    }
    case OR_FUNC_WEIERSTRASS:
        for (int j = 0; j < D; ++j) { float v = or_weier_term(z[j]); if (ACC_A(j)) A = A + v; else B = B + v; }
        return A + B;
    case OR_FUNC_SCHWEFEL:
    case OR_FUNC_REF_SCHWEFEL: { /* the constant (textbook or the reference's) from the table */
        for (int j = 0; j < D; ++j) {
            float sq = sqrtf(fabsf(z[j]));
            float tv = sq * 0.159154937f;
            float s = or_sin2pi(tv);
            float v = z[j] * s;
            if (ACC_A(j)) A = A + v; else B = B + v;
        }
        float cd = tab[0] * (float)D;
        float S = A + B;
        return cd - S;
    }
    /* ---- the reference's formulas, SG-EMTF_DE.cu:834-1069, in binary32 ---- */
    case OR_FUNC_REF_ACKLEY: { /* -20 expf(-0.2 sqrtf(S/D)) - expf(C/D) + 20 + E, C = sum cospi(2 z) */
        for (int j = 0; j < D; ++j) {
            float v = z[j] * z[j];
            float tz = 2.0f * z[j];
            float c = or_cospi(tz);
            if (ACC_A(j)) { A = A + v; A2 = A2 + c; } else { B = B + v; B2 = B2 + c; }
        }
        float Df = (float)D;
        float S = A + B, C = A2 + B2;
        float q = S / Df;
        float r = sqrtf(q);
        float a = (float)(-0.2 * (double)r); /* -0.2 is a double literal in the reference */
        float e1 = or_exp(a);
        float t1 = -20.0f * e1;
        float c2 = C / Df;
        float e2 = or_exp(c2);
        float u = t1 - e2;
        u = u + 20.0f;
        return (float)((double)u + E_REF);
    }
    case OR_FUNC_REF_RASTRIGIN: /* z^2 - 10 cospi(2 z) + 10 per dimension */
        for (int j = 0; j < D; ++j) {
            float zz = z[j] * z[j];
            float tz = 2.0f * z[j];
            float c = or_cospi(tz);
            float t = 10.0f * c;
            float v = zz - t;
            v = v + 10.0f;
            if (ACC_A(j)) A = A + v; else B = B + v;
        }
        return A + B;
    case OR_FUNC_REF_GRIEWANK:
    case OR_FUNC_REF_GRIEWANK_GA: { /* 1 + S / 4000 - prod cosf(z_j / sqrtf(j + 1)) */
        float pA = 1.0f, pB = 1.0f;
        for (int j = 0; j < D; ++j) {
            float v = z[j] * z[j];
            float q = z[j] / tab[j];
            float tq = q * 0.159154937f;
            float c = 1.0f - 2.0f * or_sinpi_sq(tq); /* cos(q) = 1 - 2 sin^2(q / 2) */
            if (ACC_A(j)) { A = A + v; pA = pA * c; } else { B = B + v; pB = pB * c; }
        }
        float P = pA * pB;
        float S = A + B;
        float s4 = S / 4000.0f;
        float u = 1.0f + s4;
        return u - P;
    }
    case OR_FUNC_REF_WEIERSTRASS: { /* sum of terms - D W_ref */
        for (int j = 0; j < D; ++j) { float v = or_weier_ref_term(z[j]); if (ACC_A(j)) A = A + v; else B = B + v; }
        float S = A + B;
        float dw = (float)D * tab[0];
        return S - dw;
    }
    case OR_FUNC_ELLIPTIC:
        for (int j = 0; j < D; ++j) {
            float zz = z[j] * z[j];
            float v = tab[j] * zz;
            if (ACC_A(j)) A = A + v; else B = B + v;
        }
        return A + B;
    default:
        return NAN;
    }
}

void or_eval32(const or_task* t, int32_t n, const float* x, float* f) {
    const int D = t->dim;
    float* w = (float*)malloc(sizeof(float) * (size_t)D * 3);
    float* z = w + D;
    float* tab = z + D;
    or_func_table(t->func, D, tab);
    for (int i = 0; i < n; ++i) {
        transform32(t, x + (size_t)i * D, w, z);
        f[i] = or_finish32(t, z, tab);
    }
    free(w);
}

/* ---- FP64 textbook forms ----------------------------------------------- */
static void transform64(const or_task* t, const double* x, double* z) {
    const int D = t->dim;
    for (int j = 0; j < D; ++j) {
        if (t->xform == OR_XFORM_ROTATE) {
            double acc = 0.0;
            for (int d = 0; d < D; ++d) acc += (double)t->rot[(size_t)j * D + d] * (x[d] - (double)t->shift[d]);
            z[j] = acc;
        } else if (t->xform == OR_XFORM_DIAG) {
            z[j] = (double)t->diag[j] * (x[j] - (double)t->shift[j]);
        } else {
            z[j] = x[j] - (double)t->shift[j];
        }
    }
}

static double func64(int func, int D, const double* z) {
    double s = 0.0, c = 0.0, p = 1.0;
    switch (func) {
    case OR_FUNC_SPHERE:
        for (int j = 0; j < D; ++j) s += z[j] * z[j];
        return s;
    case OR_FUNC_ROSENBROCK:
        for (int j = 0; j + 1 < D; ++j) {
            double a = z[j] * z[j] - z[j + 1], b = z[j] - 1.0;
            s += 100.0 * a * a + b * b;
        }
        return s;
    case OR_FUNC_ACKLEY:
        for (int j = 0; j < D; ++j) { s += z[j] * z[j]; c += cos(2.0 * M_PI * z[j]); }
        return -20.0 * exp(-0.2 * sqrt(s / D)) - exp(c / D) + 20.0 + M_E;
    case OR_FUNC_RASTRIGIN:
        for (int j = 0; j < D; ++j) s += z[j] * z[j] - 10.0 * cos(2.0 * M_PI * z[j]) + 10.0;
        return s;
    case OR_FUNC_GRIEWANK:
        for (int j = 0; j < D; ++j) { s += z[j] * z[j]; p *= cos(z[j] / sqrt((double)(j + 1))); }
        return 1.0 + s / 4000.0 - p;
    case OR_FUNC_WEIERSTRASS: {
        double w0 = 0.0;
        for (int k = 0; k <= 20; ++k) w0 += pow(0.5, k) * cos(M_PI * pow(3.0, k));
        for (int j = 0; j < D; ++j)
            for (int k = 0; k <= 20; ++k) s += pow(0.5, k) * cos(2.0 * M_PI * pow(3.0, k) * (z[j] + 0.5));
        return s - D * w0;
    }
    case OR_FUNC_SCHWEFEL:
        for (int j = 0; j < D; ++j) s += z[j] * sin(sqrt(fabs(z[j])));
        return SCHWEFEL_C * D - s;
    case OR_FUNC_ELLIPTIC:
        for (int j = 0; j < D; ++j) s += ((D > 1) ? pow(1.0e6, (double)j / (D - 1)) : 1.0) * z[j] * z[j];
        return s;
    }
    return NAN;
}

void or_eval64(const or_task* t, int32_t n, const double* x, double* f) {
    double* z = (double*)malloc(sizeof(double) * (size_t)t->dim);
    for (int i = 0; i < n; ++i) {
        transform64(t, x + (size_t)i * t->dim, z);
        f[i] = func64(t->func, t->dim, z);
    }
    free(z);
}

/* d f / d z (textbook) */
static void dfdz64(int func, int D, const double* z, double* g) {
    double s = 0.0, c = 0.0;
    for (int j = 0; j < D; ++j) g[j] = 0.0;
    switch (func) {
    case OR_FUNC_SPHERE:
        for (int j = 0; j < D; ++j) g[j] = 2.0 * z[j];
        break;
    case OR_FUNC_ROSENBROCK:
        for (int j = 0; j + 1 < D; ++j) {
            double a = z[j] * z[j] - z[j + 1];
            g[j] += 400.0 * a * z[j] + 2.0 * (z[j] - 1.0);
            g[j + 1] += -200.0 * a;
        }
        break;
    case OR_FUNC_ACKLEY: {
        for (int j = 0; j < D; ++j) { s += z[j] * z[j]; c += cos(2.0 * M_PI * z[j]); }
        double r = sqrt(s / D);
        double k1 = (r > 0.0) ? 4.0 * exp(-0.2 * r) / (r * D) : 0.0;  /* 20 * 0.2 * dr/dz */
        double k2 = exp(c / D) * 2.0 * M_PI / D;
        for (int j = 0; j < D; ++j) g[j] = k1 * z[j] + k2 * sin(2.0 * M_PI * z[j]);
        break;
    }
    case OR_FUNC_RASTRIGIN:
        for (int j = 0; j < D; ++j) g[j] = 2.0 * z[j] + 20.0 * M_PI * sin(2.0 * M_PI * z[j]);
        break;
    case OR_FUNC_GRIEWANK:
        for (int j = 0; j < D; ++j) {
            double pr = 1.0;
            for (int k = 0; k < D; ++k) if (k != j) pr *= cos(z[k] / sqrt((double)(k + 1)));
            double sj = sqrt((double)(j + 1));
            g[j] = z[j] / 2000.0 + pr * sin(z[j] / sj) / sj;
        }
        break;
    case OR_FUNC_WEIERSTRASS:
        for (int j = 0; j < D; ++j)
            for (int k = 0; k <= 20; ++k)
                g[j] -= pow(0.5, k) * 2.0 * M_PI * pow(3.0, k) * sin(2.0 * M_PI * pow(3.0, k) * (z[j] + 0.5));
        break;
    case OR_FUNC_SCHWEFEL:
        for (int j = 0; j < D; ++j) {
            double sq = sqrt(fabs(z[j]));
            g[j] = -(sin(sq) + 0.5 * sq * cos(sq));
        }
        break;
    case OR_FUNC_ELLIPTIC:
        for (int j = 0; j < D; ++j) g[j] = 2.0 * ((D > 1) ? pow(1.0e6, (double)j / (D - 1)) : 1.0) * z[j];
        break;
    }
}

void or_grad64(const or_task* t, int32_t n, const double* x, double* g) {
    const int D = t->dim;
    double* z = (double*)malloc(sizeof(double) * (size_t)D * 2);
    double* gz = z + D;
    for (int i = 0; i < n; ++i) {
        transform64(t, x + (size_t)i * D, z);
        dfdz64(t->func, D, z, gz);
        double* gx = g + (size_t)i * D;
        for (int d = 0; d < D; ++d) {
            if (t->xform == OR_XFORM_ROTATE) {
                double acc = 0.0;
                for (int j = 0; j < D; ++j) acc += (double)t->rot[(size_t)j * D + d] * gz[j];
                gx[d] = acc;
            } else if (t->xform == OR_XFORM_DIAG) {
                gx[d] = (double)t->diag[d] * gz[d];
            } else {
                gx[d] = gz[d];
            }
        }
    }
    free(z);
}


/* ---- FP32 gradient (contract arithmetic, mirrors the device adjoint) ------ */
/* d f / d z_j in binary32 from z (sequential sums and products over j). */
static void dfdz32(int func, int D, const float* z, const float* tab, float* g) {
    float s1 = 0.0f, s2 = 0.0f;
    float* pre = NULL;
    if (func == OR_FUNC_ACKLEY)
        for (int j = 0; j < D; ++j) { float v = z[j] * z[j]; s1 = s1 + v; s2 = s2 + or_sinpi_sq(z[j]); }
    if (func == OR_FUNC_GRIEWANK) {
        pre = (float*)malloc(sizeof(float) * (size_t)D);
        float p = 1.0f;
        for (int j = 0; j < D; ++j) {
            pre[j] = p;
            float t = z[j] * tab[j];
            float c = 1.0f - 2.0f * or_sinpi_sq(t);
            p = p * c;
        }
    }
    const float Df = (float)D;
    float ka = 0.0f, kb = 0.0f;
    if (func == OR_FUNC_ACKLEY) {
        float q = s1 / Df;
        float r = sqrtf(q);
        float e1 = or_expm1(-0.2f * r) + 1.0f;
        float c2 = s2 / Df;
        float e2 = or_expm1(-2.0f * c2) + 1.0f;
        float n1 = 4.0f * e1;
        float d1 = Df * r;
        ka = (r > 0.0f) ? n1 / d1 : 0.0f;
        float kc = 6.28318548f * 2.71828175f;
        kc = kc / Df;
        kb = kc * e2;
    }
    float suf = 1.0f;
    int jn = D - 1;
    /* visit j in the device order: descending slots s, j = 2 s + h for h = 0, 1 */
    for (int jj = D - 1; jj >= 0; --jj) g[jj] = 0.0f;
    for (int h = 0; h < 2; ++h) {
        suf = 1.0f;
        jn = D - 1;
        for (int s = (D + 1) / 2; s >= 0; --s) {
            const int j = 2 * s + h;
            if (j >= D) continue;
            const float zj = z[j];
            float v = 0.0f;
            switch (func) {
            case OR_FUNC_SPHERE: v = 2.0f * zj; break;
            case OR_FUNC_ELLIPTIC: { float t = 2.0f * tab[j]; v = t * zj; break; }
            case OR_FUNC_RASTRIGIN: { float a = 2.0f * zj; float b = 62.8318531f * or_sin2pi(zj); v = a + b; break; }
            case OR_FUNC_ACKLEY: { float a = ka * zj; float b = kb * or_sin2pi(zj); v = a + b; break; }
            case OR_FUNC_ROSENBROCK:
                if (j + 1 < D) {
                    float zz = zj * zj;
                    float a = zz - z[j + 1];
                    float t = 400.0f * zj;
                    t = t * a;
                    float b = zj - 1.0f;
                    b = 2.0f * b;
                    v = t + b;
                }
                if (j > 0) {
                    float zp = z[j - 1] * z[j - 1];
                    float a = zp - zj;
                    float t = 200.0f * a;
                    v = v - t;
                }
                break;
            case OR_FUNC_GRIEWANK: {
                while (jn > j) {
                    float t = z[jn] * tab[jn];
                    float c = 1.0f - 2.0f * or_sinpi_sq(t);
                    suf = suf * c;
                    --jn;
                }
                float t = zj * tab[j];
                float sn = or_sin2pi(t);
                float a = zj / 2000.0f;
                float b = pre[j] * suf;
                b = b * sn;
                float sc = 6.28318548f * tab[j];
                b = b * sc;
                v = a + b;
                break;
            }
            case OR_FUNC_WEIERSTRASS: {
                const float a = fabsf(zj);
                float acc = 0.0f;
                if (a < 8388608.0f) {
                    float c = 6.28318548f;
                    if (a < 2.32830644e-10f) {
                        float p3 = 1.0f;
                        for (int k = 0; k <= 20; ++k) {
                            float r = a * p3;
                            acc = fmaf(c, or_sin2pi(r), acc);
                            c = c * 1.5f;
                            p3 = p3 * 3.0f;
                        }
                    } else {
                        uint32_t u;
                        memcpy(&u, &a, 4);
                        const int ex = (int)((u >> 23) & 0xFF) - 127;
                        const uint64_t m = (uint64_t)((u & 0x007FFFFFu) | 0x00800000u);
                        uint64_t y = m << (64 - (23 - ex));
                        for (int k = 0; k <= 20; ++k) {
                            float r = (float)(uint32_t)(y >> 32) * 2.3283064365386963e-10f;
                            acc = fmaf(c, or_sin2pi(r), acc);
                            c = c * 1.5f;
                            y = y + (y << 1);
                        }
                    }
                }
                v = copysignf(1.0f, zj) * acc;
                break;
            }
            case OR_FUNC_SCHWEFEL: {
                float sq = sqrtf(fabsf(zj));
                float t = sq * 0.159154937f;
                float a = or_sin2pi(t);
                float hs = 0.5f * sq;
                float b = hs * or_sin2pi(t + 0.25f);
                v = -(a + b);
                break;
            }
            }
            g[j] = v;
        }
    }
    free(pre);
}

/* gx = Q^T dz: fmaf chain over j = 0 .. 2 HS - 1 (padded dimensions contribute +0) */
void or_grad32(const or_task* t, int32_t n, const float* x, float* gx) {
    const int D = t->dim;
    float* w = (float*)malloc(sizeof(float) * (size_t)D * 4);
    float* z = w + D;
    float* tab = z + D;
    float* dz = tab + D;
    or_func_table(t->func, D, tab);
    for (int i = 0; i < n; ++i) {
        transform32(t, x + (size_t)i * D, w, z);
        dfdz32(t->func, D, z, tab, dz);
        float* g = gx + (size_t)i * D;
        for (int d = 0; d < D; ++d) {
            if (t->xform == OR_XFORM_ROTATE) {
                float acc = 0.0f;
                for (int j = 0; j < D; ++j) acc = fmaf(t->rot[(size_t)j * D + d], dz[j], acc);
                g[d] = acc;
            } else if (t->xform == OR_XFORM_DIAG) {
                g[d] = fmaf(t->diag[d], dz[d], 0.0f);
            } else {
                g[d] = dz[d];
            }
        }
    }
    free(w);
}
