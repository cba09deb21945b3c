// clb_eq_device.h -- per-voxel equilibrium speciation on the device, FP64.
//
// Unknowns are the log10 activities of the nc components; species i has
// log10 c_i = logK_i + sum_j S_ij w_j, clamped to [-30, 1]
// (part4_eqsolver.hh:450-504).  Two solvers of the same mass-balance system:
//  * AA-PCF: Anderson-accelerated positive-continuous-fraction iteration,
//    depth m, Gram-Schmidt QR with condition pruning at 1e10, stop at
//    ||f||_2 < tol (part4_eqsolver.hh:593-1180);
//  * Newton on w with the analytic Jacobian ln10 sum_i S_ij S_ik c_i and a
//    one-decade step limit, stop at max_j |F_j|/T_j < tol.
//
// Register residency.  Every per-voxel array is sized by the template bounds
// NS >= ns and NC >= nc, and every loop over one runs to the bound with the
// live count as a uniform guard, fully unrolled: all array indices are then
// compile-time constants and the arrays live in VGPRs (a loop bounded by the
// runtime count puts them in scratch memory, 2.3 KB per lane, and the kernel
// becomes bound by scratch latency).  The pivot row swap is a select over the
// compile-time rows; the Anderson history and difference columns are kept
// right-aligned, so that their live window moves without a runtime index.
// The arithmetic and its order are those of the loops the guards stand for.
#pragma once

namespace clb {

constexpr int EQ_HIST = EQ_MAXD + 2;   // Anderson history slots (depth + 1, + 1 transient)

template <int NS, int NC>
__device__ __forceinline__ void eqSpecies(const EqDev &e, const double (&w)[NC], const double (&init)[NS], double (&out)[NS]) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        if (i < e.ns) {
            if (!e.isEq[i]) { out[i] = init[i] > EQ_MIN_CONC ? fmin(init[i], EQ_MAX_CONC) : EQ_MIN_CONC; continue; }
            double lc = e.logk[i];
#pragma unroll
            for (int j = 0; j < NC; ++j) if (j < e.nc) lc += e.S[i][j] * w[j];
            lc = fmax(fmin(lc, 1.0), -30.0);
            double v = exp10(lc);           // 10^lc: pow(10, lc) at half the instructions, same ulp class
            if (!isfinite(v)) v = init[i] > EQ_MIN_CONC ? init[i] : EQ_MIN_CONC;
            out[i] = fmax(fmin(v, EQ_MAX_CONC), EQ_MIN_CONC);
        }
    }
}

template <int NS, int NC>
__device__ __forceinline__ void eqTotals(const EqDev &e, const double (&c)[NS], double (&T)[NC]) {
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        if (j < e.nc) {
            double t = 0;
#pragma unroll
            for (int i = 0; i < NS; ++i)
                if (i < e.ns && e.isEq[i]) t += e.S[i][j] * fmax(fmin(c[i], EQ_MAX_CONC), EQ_MIN_CONC);
            T[j] = fmax(t, EQ_MIN_CONC);
        }
    }
}

template <int NS, int NC>
__device__ __forceinline__ void eqPcf(const EqDev &e, const double (&w)[NC], const double (&T)[NC], const double (&init)[NS],
                                      double (&f)[NC]) {
    double sp[NS];
    eqSpecies<NS, NC>(e, w, init, sp);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        if (j < e.nc) {
            double pos = 0, neg = 0;
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                if (i >= e.ns || !e.isEq[i]) continue;
                const double mu = e.S[i][j], ci = fmax(sp[i], EQ_MIN_CONC);
                if (mu > 1e-15) pos += mu * ci; else if (mu < -1e-15) neg += fabs(mu) * ci;
            }
            double SR, SP;
            if (T[j] >= 0) { SR = pos; SP = T[j] + neg; } else { SR = fabs(T[j]) + pos; SP = neg; }
            SR = fmax(SR, EQ_MIN_CONC); SP = fmax(SP, EQ_MIN_CONC);
            double v = (1.0 / e.mu0[j]) * (log10(SP) - log10(SR));
            v = fmax(fmin(v, 10.0), -10.0);
            if (!isfinite(v)) v = 0.0;
            if (T[j] <= 1e-18) v = 0.0;
            f[j] = v;
        }
    }
}

// Modified Gram-Schmidt of the m newest difference columns.  Columns are
// right-aligned: column p of DF for p in [EQ_MAXD - m, EQ_MAXD), the oldest at
// EQ_MAXD - m, processed oldest first, so dropping the oldest is m - 1.
template <int NC>
__device__ __forceinline__ double eqQR(int m, int nc, const double (&DF)[EQ_MAXD][NC], double (&Q)[EQ_MAXD][NC],
                                       double (&R)[EQ_MAXD][EQ_MAXD]) {
    const int p0 = EQ_MAXD - m;
    double rmax = 0, rmin = 1.7976931348623157e308;
#pragma unroll
    for (int a = 0; a < EQ_MAXD; ++a)
#pragma unroll
        for (int b = 0; b < EQ_MAXD; ++b) R[a][b] = 0;
#pragma unroll
    for (int p = 0; p < EQ_MAXD; ++p) {
        if (p >= p0) {
#pragma unroll
            for (int k = 0; k < NC; ++k) if (k < nc) Q[p][k] = DF[p][k];
#pragma unroll
            for (int i = 0; i < p; ++i) {
                if (i >= p0) {
                    double d = 0;
#pragma unroll
                    for (int k = 0; k < NC; ++k) if (k < nc) d += Q[i][k] * DF[p][k];
                    R[i][p] = d;
#pragma unroll
                    for (int k = 0; k < NC; ++k) if (k < nc) Q[p][k] -= d * Q[i][k];
                }
            }
            double s = 0;
#pragma unroll
            for (int k = 0; k < NC; ++k) if (k < nc) s += Q[p][k] * Q[p][k];
            R[p][p] = sqrt(s);
            if (R[p][p] > 1e-15) {
#pragma unroll
                for (int k = 0; k < NC; ++k) if (k < nc) Q[p][k] /= R[p][p];
            }
            const double a = fabs(R[p][p]);
            if (a > rmax) rmax = a;
            if (a > 1e-30 && a < rmin) rmin = a;
        }
    }
    return (rmin > 1e-30) ? rmax / rmin : 1.7976931348623157e308;
}

// The history is right-aligned too: hn live entries in slots [EQ_HIST - hn,
// EQ_HIST), the newest last; a new entry shifts the slots down by one, and
// forgetting the oldest is hn - 1.
template <int NS, int NC>
__device__ __forceinline__ bool eqAnderson(const EqDev &e, const double (&init)[NS], const double (&T)[NC], double (&out)[NS]) {
    const int nc = e.nc;
    const int depth = e.depth < 1 ? 1 : (e.depth > EQ_MAXD ? EQ_MAXD : e.depth);
    double om[NC], onew[NC], fk[NC];
    double oh[EQ_HIST][NC], fh[EQ_HIST][NC];
    int hn = 0;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        if (j < nc) {
            double cc = 1e-7;
            const int si = e.compSpecies[j];
#pragma unroll
            for (int s = 0; s < NS; ++s) if (s == si && s < e.ns) cc = init[s];
            cc = fmax(fmin(cc, EQ_MAX_CONC), EQ_MIN_CONC);
            om[j] = log10(cc);
        }
    }
    eqPcf<NS, NC>(e, om, T, init, fk);
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        if (j < nc) {
            onew[j] = fmax(fmin(om[j] + fk[j], 1.0), -30.0);
            oh[EQ_HIST - 1][j] = om[j];
            fh[EQ_HIST - 1][j] = fk[j];
        }
    }
    hn = 1;
#pragma unroll
    for (int j = 0; j < NC; ++j) if (j < nc) om[j] = onew[j];
    for (int it = 1; it < e.maxIter; ++it) {
        eqPcf<NS, NC>(e, om, T, init, fk);
        double fn = 0;
#pragma unroll
        for (int j = 0; j < NC; ++j) if (j < nc) fn += fk[j] * fk[j];
        fn = sqrt(fn);
        if (fn < e.tol) { eqSpecies<NS, NC>(e, om, init, out); return true; }
#pragma unroll
        for (int h = 0; h + 1 < EQ_HIST; ++h)
#pragma unroll
            for (int j = 0; j < NC; ++j) { oh[h][j] = oh[h + 1][j]; fh[h][j] = fh[h + 1][j]; }
#pragma unroll
        for (int j = 0; j < NC; ++j) { oh[EQ_HIST - 1][j] = om[j]; fh[EQ_HIST - 1][j] = fk[j]; }
        ++hn;
        const int mk = depth < hn - 1 ? depth : hn - 1;
        if (mk < 1) {
#pragma unroll
            for (int j = 0; j < NC; ++j) if (j < nc) onew[j] = om[j] + fk[j];
        } else {
            // difference column p is entry (slot q) minus entry (slot q - 1),
            // q = EQ_HIST - EQ_MAXD + p: the newest difference sits at p = EQ_MAXD - 1
            double DF[EQ_MAXD][NC], DX[EQ_MAXD][NC], Q[EQ_MAXD][NC], R[EQ_MAXD][EQ_MAXD];
#pragma unroll
            for (int p = 0; p < EQ_MAXD; ++p) {
                if (p >= EQ_MAXD - mk) {
                    constexpr int dq = EQ_HIST - EQ_MAXD;
#pragma unroll
                    for (int j = 0; j < NC; ++j) {
                        if (j < nc) {
                            DF[p][j] = fh[dq + p][j] - fh[dq + p - 1][j];
                            DX[p][j] = oh[dq + p][j] - oh[dq + p - 1][j];
                        }
                    }
                }
            }
            int m = mk;
            double cond = eqQR<NC>(m, nc, DF, Q, R);
            while (cond > 1e10 && m > 1) {       // drop the oldest difference
                --m;
                cond = eqQR<NC>(m, nc, DF, Q, R);
            }
            const int p0 = EQ_MAXD - m;
            double qtf[EQ_MAXD], gam[EQ_MAXD];
#pragma unroll
            for (int p = 0; p < EQ_MAXD; ++p) {
                if (p >= p0) {
                    double d = 0;
#pragma unroll
                    for (int j = 0; j < NC; ++j) if (j < nc) d += Q[p][j] * fk[j];
                    qtf[p] = d;
                }
            }
#pragma unroll
            for (int p = EQ_MAXD - 1; p >= 0; --p) {
                if (p >= p0) {
                    gam[p] = qtf[p];
#pragma unroll
                    for (int q = p + 1; q < EQ_MAXD; ++q) gam[p] -= R[p][q] * gam[q];
                    if (fabs(R[p][p]) > 1e-30) gam[p] /= R[p][p];
                }
            }
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                if (j < nc) {
                    double dx = 0, df = 0;
#pragma unroll
                    for (int p = 0; p < EQ_MAXD; ++p) if (p >= p0) { dx += DX[p][j] * gam[p]; df += DF[p][j] * gam[p]; }
                    onew[j] = om[j] - dx + 1.0 * (fk[j] - df);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            if (j < nc) {
                onew[j] = fmax(fmin(onew[j], 1.0), -30.0);
                if (!isfinite(onew[j])) onew[j] = om[j];
            }
        }
#pragma unroll
        for (int j = 0; j < NC; ++j) if (j < nc) om[j] = onew[j];
        if (hn > depth + 1) hn = depth + 1;     // forget the oldest entries
    }
    eqSpecies<NS, NC>(e, om, init, out);
    return false;
}

template <int NS, int NC>
__device__ __forceinline__ bool eqNewton(const EqDev &e, const double (&init)[NS], const double (&T)[NC], double (&out)[NS]) {
    const int nc = e.nc;
    double om[NC], sp[NS];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        if (j < nc) {
            double cc = 1e-7;
            const int si = e.compSpecies[j];
#pragma unroll
            for (int s = 0; s < NS; ++s) if (s == si && s < e.ns) cc = init[s];
            cc = fmax(fmin(cc, EQ_MAX_CONC), EQ_MIN_CONC);
            om[j] = log10(cc);
        }
    }
    const double LN10 = 2.302585092994045684;
    bool conv = false;
    for (int it = 0; it < e.maxIter; ++it) {
        eqSpecies<NS, NC>(e, om, init, sp);
        // the Jacobian and the right-hand side -F
        double A[NC][NC], b[NC];
        double rmax = 0;
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            if (j < nc) {
                double t = 0;
#pragma unroll
                for (int i = 0; i < NS; ++i) if (i < e.ns && e.isEq[i]) t += e.S[i][j] * sp[i];
                const double F = t - T[j];
                double r = fabs(F) / fmax(fabs(T[j]), 1e-30);
                if (T[j] <= 1e-18) r = 0;
                rmax = fmax(rmax, r);
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                    if (k < nc) {
                        double a = 0;
#pragma unroll
                        for (int i = 0; i < NS; ++i) if (i < e.ns && e.isEq[i]) a += e.S[i][j] * e.S[i][k] * sp[i];
                        A[j][k] = LN10 * a;
                    }
                }
                b[j] = -F;
                if (T[j] <= 1e-18) {
#pragma unroll
                    for (int k = 0; k < NC; ++k) if (k < nc) A[j][k] = (k == j) ? 1.0 : 0.0;
                    b[j] = 0.0;
                }
            }
        }
        if (rmax < e.tol) { conv = true; break; }
        // Gaussian elimination with partial pivoting (first largest |A[r][k]|)
        bool sing = false;
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            int pv = k;
            if (k < nc && !sing) {
                double bv = fabs(A[k][k]);
#pragma unroll
                for (int r = k + 1; r < NC; ++r) if (r < nc && fabs(A[r][k]) > bv) { pv = r; bv = fabs(A[r][k]); }
                if (bv < 1e-300) sing = true;    // the elimination stops at the first singular column
            }
            if (k < nc && !sing) {
                // columns left of k are never read again: the swap covers k..nc-1 and b
#pragma unroll
                for (int r = k + 1; r < NC; ++r) {
                    if (r < nc) {
                        const bool sw = r == pv;
#pragma unroll
                        for (int c = k; c < NC; ++c) {
                            if (c < nc) {
                                const double t = A[k][c];
                                A[k][c] = sw ? A[r][c] : t;
                                A[r][c] = sw ? t : A[r][c];
                            }
                        }
                        const double t = b[k];
                        b[k] = sw ? b[r] : t;
                        b[r] = sw ? t : b[r];
                    }
                }
#pragma unroll
                for (int r = k + 1; r < NC; ++r) {
                    if (r < nc) {
                        const double fct = A[r][k] / A[k][k];
#pragma unroll
                        for (int c = k; c < NC; ++c) if (c < nc) A[r][c] -= fct * A[k][c];
                        b[r] -= fct * b[k];
                    }
                }
            }
        }
        if (sing) break;
        double d[NC];
#pragma unroll
        for (int k = NC - 1; k >= 0; --k) {
            if (k < nc) {
                double s = b[k];
#pragma unroll
                for (int c = k + 1; c < NC; ++c) if (c < nc) s -= A[k][c] * d[c];
                d[k] = s / A[k][k];
            }
        }
        double dmax = 0;
#pragma unroll
        for (int k = 0; k < NC; ++k) if (k < nc) dmax = fmax(dmax, fabs(d[k]));
        const double sc = dmax > 1.0 ? 1.0 / dmax : 1.0;
#pragma unroll
        for (int k = 0; k < NC; ++k) if (k < nc) om[k] = fmax(fmin(om[k] + sc * d[k], 1.0), -30.0);
    }
    eqSpecies<NS, NC>(e, om, init, out);
    return conv;
}

// calculate_species_concentrations (part4_eqsolver.hh:1195-1222); SOLVER is
// e.solver (EQ_SOLVER_AAPCF or EQ_SOLVER_NEWTON), one solver per kernel
template <int NS, int NC, int SOLVER>
__device__ __forceinline__ bool eqSolve(const EqDev &e, const double (&c)[NS], double (&out)[NS]) {
    if (e.nc == 0 || e.ns == 0) {
#pragma unroll
        for (int i = 0; i < NS; ++i) if (i < e.ns) out[i] = c[i];
        return true;
    }
    double T[NC];
    eqTotals<NS, NC>(e, c, T);
    const bool conv = SOLVER == EQ_SOLVER_NEWTON ? eqNewton<NS, NC>(e, c, T, out) : eqAnderson<NS, NC>(e, c, T, out);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
        if (i < e.ns) {
            if (!isfinite(out[i])) out[i] = c[i] > EQ_MIN_CONC ? fmin(c[i], EQ_MAX_CONC) : EQ_MIN_CONC;
            out[i] = fmax(fmin(out[i], EQ_MAX_CONC), EQ_MIN_CONC);
        }
    }
    return conv;
}

}  // namespace clb
