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
// Everything lives in registers / scratch of fixed size (nc <= 8, m <= 4).
#pragma once

namespace clb {

__device__ __forceinline__ void eqSpecies(const EqDev &e, const double *w, const double *init, double *out) {
    for (int i = 0; i < e.ns; ++i) {
        if (!e.isEq[i]) { out[i] = init[i] > EQ_MIN_CONC ? fmin(init[i], EQ_MAX_CONC) : EQ_MIN_CONC; continue; }
        double lc = e.logk[i];
        for (int j = 0; j < e.nc; ++j) lc += e.S[i][j] * w[j];
        lc = fmax(fmin(lc, 1.0), -30.0);
        double v = pow(10.0, lc);
        if (!isfinite(v)) v = init[i] > EQ_MIN_CONC ? init[i] : EQ_MIN_CONC;
        out[i] = fmax(fmin(v, EQ_MAX_CONC), EQ_MIN_CONC);
    }
}

__device__ __forceinline__ void eqTotals(const EqDev &e, const double *c, double *T) {
    for (int j = 0; j < e.nc; ++j) {
        double t = 0;
        for (int i = 0; i < e.ns; ++i) if (e.isEq[i]) t += e.S[i][j] * fmax(fmin(c[i], EQ_MAX_CONC), EQ_MIN_CONC);
        T[j] = fmax(t, EQ_MIN_CONC);
    }
}

__device__ __forceinline__ void eqPcf(const EqDev &e, const double *w, const double *T, const double *init, double *f) {
    double sp[EQ_MAXS];
    eqSpecies(e, w, init, sp);
    for (int j = 0; j < e.nc; ++j) {
        double pos = 0, neg = 0;
        for (int i = 0; i < e.ns; ++i) {
            if (!e.isEq[i]) continue;
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

// modified Gram-Schmidt of the m columns DF[0..m-1] (each nc long)
__device__ __forceinline__ double eqQR(int m, int nc, double DF[EQ_MAXD][EQ_MAXC], double Q[EQ_MAXD][EQ_MAXC],
                                       double R[EQ_MAXD][EQ_MAXD]) {
    double rmax = 0, rmin = 1.7976931348623157e308;
    for (int a = 0; a < m; ++a) for (int b = 0; b < m; ++b) R[a][b] = 0;
    for (int j = 0; j < m; ++j) {
        for (int k = 0; k < nc; ++k) Q[j][k] = DF[j][k];
        for (int i = 0; i < j; ++i) {
            double d = 0;
            for (int k = 0; k < nc; ++k) d += Q[i][k] * DF[j][k];
            R[i][j] = d;
            for (int k = 0; k < nc; ++k) Q[j][k] -= d * Q[i][k];
        }
        double s = 0;
        for (int k = 0; k < nc; ++k) s += Q[j][k] * Q[j][k];
        R[j][j] = sqrt(s);
        if (R[j][j] > 1e-15) for (int k = 0; k < nc; ++k) Q[j][k] /= R[j][j];
        const double a = fabs(R[j][j]);
        if (a > rmax) rmax = a;
        if (a > 1e-30 && a < rmin) rmin = a;
    }
    return (rmin > 1e-30) ? rmax / rmin : 1.7976931348623157e308;
}

__device__ bool eqAnderson(const EqDev &e, const double *init, const double *T, double *out) {
    const int nc = e.nc;
    const int depth = e.depth < 1 ? 1 : (e.depth > EQ_MAXD ? EQ_MAXD : e.depth);
    double om[EQ_MAXC], onew[EQ_MAXC], fk[EQ_MAXC];
    // history ring of at most depth+1 entries (+1 transient)
    double oh[EQ_MAXD + 2][EQ_MAXC], fh[EQ_MAXD + 2][EQ_MAXC];
    int hn = 0;
    for (int j = 0; j < nc; ++j) {
        double cc = 1e-7;
        const int si = e.compSpecies[j];
        if (si >= 0 && si < e.ns) cc = init[si];
        cc = fmax(fmin(cc, EQ_MAX_CONC), EQ_MIN_CONC);
        om[j] = log10(cc);
    }
    eqPcf(e, om, T, init, fk);
    for (int j = 0; j < nc; ++j) { onew[j] = fmax(fmin(om[j] + fk[j], 1.0), -30.0); oh[0][j] = om[j]; fh[0][j] = fk[j]; }
    hn = 1;
    for (int j = 0; j < nc; ++j) om[j] = onew[j];
    for (int it = 1; it < e.maxIter; ++it) {
        eqPcf(e, om, T, init, fk);
        double fn = 0;
        for (int j = 0; j < nc; ++j) fn += fk[j] * fk[j];
        fn = sqrt(fn);
        if (fn < e.tol) { eqSpecies(e, om, init, out); return true; }
        for (int j = 0; j < nc; ++j) { oh[hn][j] = om[j]; fh[hn][j] = fk[j]; }
        ++hn;
        int mk = depth < hn - 1 ? depth : hn - 1;
        if (mk < 1) {
            for (int j = 0; j < nc; ++j) onew[j] = om[j] + fk[j];
        } else {
            double DF[EQ_MAXD][EQ_MAXC], DX[EQ_MAXD][EQ_MAXC], Q[EQ_MAXD][EQ_MAXC], R[EQ_MAXD][EQ_MAXD];
            for (int i = 0; i < mk; ++i) {
                const int ix = hn - mk - 1 + i;
                for (int j = 0; j < nc; ++j) { DF[i][j] = fh[ix + 1][j] - fh[ix][j]; DX[i][j] = oh[ix + 1][j] - oh[ix][j]; }
            }
            int m = mk;
            double cond = eqQR(m, nc, DF, Q, R);
            while (cond > 1e10 && m > 1) {
                for (int i = 0; i + 1 < m; ++i) for (int j = 0; j < nc; ++j) { DF[i][j] = DF[i + 1][j]; DX[i][j] = DX[i + 1][j]; }
                --m;
                cond = eqQR(m, nc, DF, Q, R);
            }
            double qtf[EQ_MAXD], gam[EQ_MAXD];
            for (int i = 0; i < m; ++i) { double d = 0; for (int j = 0; j < nc; ++j) d += Q[i][j] * fk[j]; qtf[i] = d; }
            for (int i = m - 1; i >= 0; --i) {
                gam[i] = qtf[i];
                for (int j = i + 1; j < m; ++j) gam[i] -= R[i][j] * gam[j];
                if (fabs(R[i][i]) > 1e-30) gam[i] /= R[i][i];
            }
            for (int j = 0; j < nc; ++j) {
                double dx = 0, df = 0;
                for (int i = 0; i < m; ++i) { dx += DX[i][j] * gam[i]; df += DF[i][j] * gam[i]; }
                onew[j] = om[j] - dx + 1.0 * (fk[j] - df);
            }
        }
        for (int j = 0; j < nc; ++j) {
            onew[j] = fmax(fmin(onew[j], 1.0), -30.0);
            if (!isfinite(onew[j])) onew[j] = om[j];
        }
        for (int j = 0; j < nc; ++j) om[j] = onew[j];
        while (hn > depth + 1) {
            for (int i = 0; i + 1 < hn; ++i) for (int j = 0; j < nc; ++j) { oh[i][j] = oh[i + 1][j]; fh[i][j] = fh[i + 1][j]; }
            --hn;
        }
    }
    eqSpecies(e, om, init, out);
    return false;
}

__device__ bool eqNewton(const EqDev &e, const double *init, const double *T, double *out) {
    const int nc = e.nc;
    double om[EQ_MAXC], sp[EQ_MAXS];
    for (int j = 0; j < nc; ++j) {
        double cc = 1e-7;
        const int si = e.compSpecies[j];
        if (si >= 0 && si < e.ns) cc = init[si];
        cc = fmax(fmin(cc, EQ_MAX_CONC), EQ_MIN_CONC);
        om[j] = log10(cc);
    }
    const double LN10 = 2.302585092994045684;
    bool conv = false;
    for (int it = 0; it < e.maxIter; ++it) {
        eqSpecies(e, om, init, sp);
        double A[EQ_MAXC][EQ_MAXC + 1];
        double rmax = 0;
        for (int j = 0; j < nc; ++j) {
            double t = 0;
            for (int i = 0; i < e.ns; ++i) if (e.isEq[i]) t += e.S[i][j] * sp[i];
            const double F = t - T[j];
            double r = fabs(F) / fmax(fabs(T[j]), 1e-30);
            if (T[j] <= 1e-18) r = 0;
            rmax = fmax(rmax, r);
            for (int k = 0; k < nc; ++k) {
                double a = 0;
                for (int i = 0; i < e.ns; ++i) if (e.isEq[i]) a += e.S[i][j] * e.S[i][k] * sp[i];
                A[j][k] = LN10 * a;
            }
            A[j][nc] = -F;
            if (T[j] <= 1e-18) { for (int k = 0; k < nc; ++k) A[j][k] = (k == j) ? 1.0 : 0.0; A[j][nc] = 0.0; }
        }
        if (rmax < e.tol) { conv = true; break; }
        bool sing = false;
        for (int k = 0; k < nc; ++k) {
            int pv = k;
            for (int r = k + 1; r < nc; ++r) if (fabs(A[r][k]) > fabs(A[pv][k])) pv = r;
            if (fabs(A[pv][k]) < 1e-300) { sing = true; break; }
            if (pv != k) for (int c = 0; c <= nc; ++c) { const double t = A[k][c]; A[k][c] = A[pv][c]; A[pv][c] = t; }
            for (int r = k + 1; r < nc; ++r) {
                const double fct = A[r][k] / A[k][k];
                for (int c = k; c <= nc; ++c) A[r][c] -= fct * A[k][c];
            }
        }
        if (sing) break;
        double d[EQ_MAXC];
        for (int k = nc - 1; k >= 0; --k) { double s = A[k][nc]; for (int c = k + 1; c < nc; ++c) s -= A[k][c] * d[c]; d[k] = s / A[k][k]; }
        double dmax = 0;
        for (int k = 0; k < nc; ++k) dmax = fmax(dmax, fabs(d[k]));
        const double sc = dmax > 1.0 ? 1.0 / dmax : 1.0;
        for (int k = 0; k < nc; ++k) om[k] = fmax(fmin(om[k] + sc * d[k], 1.0), -30.0);
    }
    eqSpecies(e, om, init, out);
    return conv;
}

// calculate_species_concentrations (part4_eqsolver.hh:1195-1222)
__device__ bool eqSolve(const EqDev &e, const double *c, double *out) {
    if (e.nc == 0 || e.ns == 0) { for (int i = 0; i < e.ns; ++i) out[i] = c[i]; return true; }
    double T[EQ_MAXC];
    eqTotals(e, c, T);
    const bool conv = e.solver == 1 ? eqNewton(e, c, T, out) : eqAnderson(e, c, T, out);
    for (int i = 0; i < e.ns; ++i) {
        if (!isfinite(out[i])) out[i] = c[i] > EQ_MIN_CONC ? fmin(c[i], EQ_MAX_CONC) : EQ_MIN_CONC;
        out[i] = fmax(fmin(out[i], EQ_MAX_CONC), EQ_MIN_CONC);
    }
    return conv;
}

}  // namespace clb
