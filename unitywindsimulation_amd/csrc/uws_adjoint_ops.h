// Reverse by hand of the corrections, the PISO coefficients and aodu, and the
// Rhie-Chow coefficients, in gather form.  Included by uws_kernels.hip inside
// its anonymous namespace, after the forward bodies and helpers it uses.
//
// Every kernel runs one thread per cell R and accumulates only the adjoints
// of R's inputs: R's own share and the shares of the neighbours whose outputs
// read R, recomputed from those neighbours' output adjoints and records.  The
// forward bodies differentiated here are body_vel_correct, body_piso_correct,
// body_piso_precompute, body_piso_aodu and body_pres_precompute (FvmColl.compute
// and PisoColl.compute, cited there); a cell's outputs depend on its inputs
// linearly except through the mean 1/D of a face, 1/D itself and 1/presCoeff.
#pragma once

// index of the cell k steps from (x, y, z) in direction d, -1 outside the grid
template <class A>
__device__ __forceinline__ int step_of(const ParamsT<Dual<A>>& p, int x, int y, int z, int d, int k) {
    const int nx = x + k * c_off[d][0], ny = y + k * c_off[d][1], nz = z + k * c_off[d][2];
    return inreg(p, nx, ny, nz) ? lin(p, nx, ny, nz) : -1;
}
__device__ __forceinline__ bool fluid_at(const uint8_t* flag, int i) { return i >= 0 && flag[i] == UWS_CELL_FLUID; }

// ---------------------------------------------------------------------------
// Velocity corrections (body_vel_correct, body_piso_correct with
// K_NO_INPLACE).  u'_c(X) depends on the face values pf(X, +c) - pf(X, -c)
// with the factor k_c(X); a face value averages p' of X and a fluid
// neighbour, or is p'(X) alone (or 0 at a fixed-pressure domain face, SIMPLE).
// ---------------------------------------------------------------------------
template <class T, class A>
__device__ __forceinline__ A corr_k(const AdjCorrect<T, A>& a, A dvden, A dx, int X, int c) {
    const A lam = a.lam_uc[c][X];
    const A v = (A)a.coef[c][X];
    // SIMPLE: -(dv/den) (1/D) / dx;  PISO: -(1/D) (dv/den) / dx
    return a.piso ? -v * dvden / dx * lam : -dvden * ((A)1 / v) / dx * lam;
}

template <class T, class A>
__global__ __launch_bounds__(256) void k_adj_correct(ParamsT<Dual<A>> p, AdjCorrect<T, A> a) {
    const int R = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (R >= p.N || !fluid_at(a.flag, R)) return;
    const int x = R % p.X, t = R / p.X, y = t % p.Y, z = t / p.Y;
    const A dvden = p.dv.v / p.den.v, dx = p.dx.v;
    int Q[6];
    bool fl[6];
    const A pcR = (A)a.pc[R];
    A pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        Q[d] = step_of(p, x, y, z, d, 1);
        fl[d] = fluid_at(a.flag, Q[d]);
        const bool zero_bnd = !a.piso && Q[d] < 0 && p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE;
        pf[d] = fl[d] ? (A)0.5 * (pcR + (A)a.pc[Q[d]]) : (zero_bnd ? (A)0 : pcR);
    }
    A lpc = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const A lam = a.lam_uc[c][R];
        const A diff = pf[2 * c] - pf[2 * c + 1];
        if (a.piso) {
            a.lam_coef[c][R] += -((A)a.ao[c][R] + dvden * (diff / dx)) * lam;
            a.lam_ao[c][R] += -(A)a.coef[c][R] * lam;
        } else {
            const A D = (A)a.coef[c][R];
            a.lam_coef[c][R] += dvden * diff / (dx * D * D) * lam;
        }
        const A k = corr_k(a, dvden, dx, R, c);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int d = 2 * c + s;
            const A lpf = s ? -k : k;
            const bool zero_bnd = !a.piso && Q[d] < 0 && p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE;
            lpc += fl[d] ? (A)0.5 * lpf : (zero_bnd ? (A)0 : lpf);
        }
    }
    // a fluid neighbour's face towards R averages p'(R) in
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!fl[d]) continue;
        const A k = corr_k(a, dvden, dx, Q[d], d >> 1);
        lpc += (A)0.5 * ((d & 1) ? k : -k);  // the face is the neighbour's +c face when d is a - direction
    }
    a.lam_pc[R] += lpc;
}

// ---------------------------------------------------------------------------
// PISO coefficients (body_piso_precompute): presCoeff = -sum of the fluid
// faces' coefficients (each stored on the cell owning the face: the + faces
// on the cell, the - faces on the neighbour), velTerm = sum over the fluid
// faces of -+ AdInv_f AodU_f with the face means of 1/D and aodu.
// ---------------------------------------------------------------------------
template <class T, class A>
__global__ __launch_bounds__(256) void k_adj_piso_lcf(ParamsT<Dual<A>> p, AdjPisoCoef<T, A> a) {
    const int R = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (R >= p.N) return;
    A l = 0;
    if (fluid_at(a.flag, R)) {
        const int x = R % p.X, t = R / p.X, y = t % p.Y, z = t / p.Y;
        A pc = 0;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int Q = step_of(p, x, y, z, d, 1);
            if (!fluid_at(a.flag, Q)) continue;
            const int c = d >> 1, owner = (d & 1) == 0 ? R : Q;
            pc -= (A)(c == 0 ? a.cpe[owner] : a.cp[c][owner]);
        }
        // d(1/presCoeff)/d(face coefficient) = 1/presCoeff^2
        l = a.lam_idp[R] / (pc * pc);
    }
    a.lcf[R] = l;
}

template <class T, class A>
__global__ __launch_bounds__(256) void k_adj_piso_coef(ParamsT<Dual<A>> p, AdjPisoCoef<T, A> a) {
    const int R = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (R >= p.N || !fluid_at(a.flag, R)) return;
    const int x = R % p.X, t = R / p.X, y = t % p.Y, z = t / p.Y;
    const A lbR = a.lam_bp[R], lcfR = a.lcf[R];
    A linv[3] = {0, 0, 0}, lao[3] = {0, 0, 0}, lcoef[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int Q = step_of(p, x, y, z, d, 1);
        if (!fluid_at(a.flag, Q)) continue;
        const int c = d >> 1;
        const bool plus = (d & 1) == 0;
        const A AdInv = (A)0.5 * ((A)a.invD[c][R] + (A)a.invD[c][Q]);
        const A AodU = (A)0.5 * ((A)a.ao[c][R] + (A)a.ao[c][Q]);
        // R's face d (velTerm -= on + faces, += on - faces) and Q's face
        // back towards R (the opposite sign), both averaging R's values
        const A sR = plus ? (A)-1 : (A)1;
        const A lb = sR * (lbR - a.lam_bp[Q]);
        linv[c] += (A)0.5 * lb * AodU;
        lao[c] += (A)0.5 * lb * AdInv;
        // the coefficient of a + face of R is R's: read by R and by Q
        if (plus) lcoef[c] += lcfR + a.lcf[Q];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.lam_invD[c][R] += linv[c];
        a.lam_ao[c][R] += lao[c];
    }
    a.lam_cpe[R] += lcoef[0];
    a.lam_cp[1][R] += lcoef[1];
    a.lam_cp[2][R] += lcoef[2];
}

// ---------------------------------------------------------------------------
// PISO aodu (body_piso_aodu): for a fluid cell P and component i,
// aodu_i(P) = sum over fluid faces f of (+-q uc_i(Q_f) (u_c(P) + u_c(Q_f)) - g uc_i(Q_f)),
// q = ds/4, g = nu ds/dx, c the face's axis.
// ---------------------------------------------------------------------------
template <class T, class A>
__global__ __launch_bounds__(256) void k_adj_aodu(ParamsT<Dual<A>> p, AdjAodu<T, A> a) {
    const int R = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (R >= p.N || !fluid_at(a.flag, R)) return;
    const int x = R % p.X, t = R / p.X, y = t % p.Y, z = t / p.Y;
    const A q = p.ds.v * (A)0.25, g = p.nu.v * p.ds.v / p.dx.v;
    const A lR[3] = {a.lam_ao[0][R], a.lam_ao[1][R], a.lam_ao[2][R]};
    const A ucR[3] = {(A)a.uc[0][R], (A)a.uc[1][R], (A)a.uc[2][R]};
    A lu[3] = {0, 0, 0}, luc[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int Q = step_of(p, x, y, z, d, 1);
        if (!fluid_at(a.flag, Q)) continue;
        const int c = d >> 1;
        const A sp = (d & 1) == 0 ? (A)1 : (A)-1;
        const A lQ[3] = {a.lam_ao[0][Q], a.lam_ao[1][Q], a.lam_ao[2][Q]};
        // R's own face: both u_c(R) and u_c(Q) enter with sp q uc_i(Q)
        lu[c] += sp * q * (lR[0] * (A)a.uc[0][Q] + lR[1] * (A)a.uc[1][Q] + lR[2] * (A)a.uc[2][Q]);
        // Q's face towards R (sign -sp) reads u_c(R) and uc_i(R)
        lu[c] -= sp * q * (lQ[0] * ucR[0] + lQ[1] * ucR[1] + lQ[2] * ucR[2]);
        const A sum = (A)a.u[c][Q] + (A)a.u[c][R];
#pragma unroll
        for (int i = 0; i < 3; ++i) luc[i] += lQ[i] * (-sp * q * sum - g);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.lam_u[c][R] += lu[c];
        a.lam_uc[c][R] += luc[c];
    }
}

// ---------------------------------------------------------------------------
// Rhie-Chow coefficients (body_pres_precompute; the Dual runs use the
// sequential form body_pres_precompute_seq, whose fixed-pressure boundary
// term reads 1/D as 1 / D).  For a fluid cell P and a fluid face f with axis
// c and sign s (+1 on E/N/U):
//   Dp -= (dv/den) DInv_f / dx,   DInv_f = (1/D_c(P) + 1/D_c(Q)) / 2,
//   bp += s ((u_c(P) + u_c(Q)) / 2 - (dv DInv_f / den) (g - ga)),
// where (g - ga) = s/dx (cQ p_Q + cP p_P + cF p_F + cB p_B): p_F two cells
// out, p_B the cell behind P, the weights set by which of F, B are fluid:
//   F and B: (3/4, -3/4, -1/4, 1/4)   F only: (3/4, -1/2, -1/4, 0)
//   B only:  (1/2, -3/4, 0, 1/4)      neither: 0 (ga = g).
// Domain faces add s u_bnd (fixed value) or s u_c(P) (zero gradient) to bp
// and -2 (dv/den) (1 / D_c(P)) / dx to Dp at a fixed-pressure face.  The
// owned + faces also give cp_c(P) = cpE(P) = (dv/den) DInv_f / dx.
// k_adj_rc_local writes each fluid cell's face adjoints (w, ldi) and its own
// shares of u, p, D and the textures; k_adj_rc_gather adds the neighbours'
// shares and the mean-1/D adjoints of the faces on both sides.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rc_weights(bool fF, bool bF, double& cQ, double& cP, double& cF, double& cB) {
    cF = cB = 0.0;
    if (fF && bF) { cQ = 0.75; cP = -0.75; cF = -0.25; cB = 0.25; }
    else if (fF) { cQ = 0.75; cP = -0.5; cF = -0.25; }
    else if (bF) { cQ = 0.5; cP = -0.75; cB = 0.25; }
    else { cQ = 0.0; cP = 0.0; }
}

template <class T, class A>
__global__ __launch_bounds__(256) void k_adj_rc_local(ParamsT<Dual<A>> p, AdjRhieChow<T, A> a) {
    const int P = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (P >= p.N) return;
    A w[6] = {0, 0, 0, 0, 0, 0}, ldi[6] = {0, 0, 0, 0, 0, 0};
    if (fluid_at(a.flag, P)) {
        const int x = P % p.X, t = P / p.X, y = t % p.Y, z = t / p.Y;
        const A dv = p.dv.v, den = p.den.v, dvden = dv / den, dx = p.dx.v;
        int Q[6];
        bool fl[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            Q[d] = step_of(p, x, y, z, d, 1);
            fl[d] = fluid_at(a.flag, Q[d]);
        }
        A Dp = 0;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int c = d >> 1;
            if (fl[d]) Dp -= dvden * ((A)0.5 * ((A)a.invD[c][P] + (A)a.invD[c][Q[d]])) / dx;
            else if (Q[d] < 0 && p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE)
                Dp -= (A)2 * dvden * ((A)1 / (A)a.D[c][P]) / dx;
        }
        const A lDp = -a.lam_idp[P] / (Dp * Dp);
        const A lbp = a.lam_bp[P];
        const A pP = (A)a.p[P];
        A lu[3] = {0, 0, 0}, lD[3] = {0, 0, 0}, lp = 0;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int c = d >> 1;
            const A sg = (d & 1) == 0 ? (A)1 : (A)-1;
            if (fl[d]) {
                const int F = step_of(p, x, y, z, d, 2);
                const bool fF = fluid_at(a.flag, F), bF = fl[d ^ 1];
                double cQ, cP, cF, cB;
                rc_weights(fF, bF, cQ, cP, cF, cB);
                A gga = (A)cQ * (A)a.p[Q[d]] + (A)cP * pP;
                if (fF) gga += (A)cF * (A)a.p[F];
                if (bF) gga += (A)cB * (A)a.p[Q[d ^ 1]];
                gga *= sg / dx;
                const A DInv = (A)0.5 * ((A)a.invD[c][P] + (A)a.invD[c][Q[d]]);
                ldi[d] = -dvden / dx * lDp - sg * dv / den * gga * lbp;
                w[d] = sg * (-sg * dv * DInv / den * lbp) / dx;
                lu[c] += (A)0.5 * sg * lbp;
                lp += (A)cP * w[d];
            } else if (Q[d] < 0) {
                const int bf = bface_of(d);
                const int tv = p.velBnd[bf];
                if (tv == UWS_VEL_FIXED_VALUE) {
                    if (a.lam_bnd[bf][c]) a.lam_bnd[bf][c][bnd_elem(p, d, x, y, z)] += sg * lbp;
                } else if (tv == UWS_VEL_ZERO_GRAD) {
                    lu[c] += sg * lbp;
                }
                if (p.presBnd[bf] == UWS_PRES_FIXED_VALUE) {
                    const A Dc = (A)a.D[c][P];
                    lD[c] += lDp * (A)2 * dvden / (dx * Dc * Dc);
                }
            }
        }
        // the owned + faces' coefficients cp_c and cpE
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (fl[2 * c]) ldi[2 * c] += dvden / dx * (a.lam_cp[c][P] + (c == 0 ? a.lam_cpe[P] : (A)0));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a.lam_u[c][P] += lu[c];
            a.lam_D[c][P] += lD[c];
        }
        a.lam_p[P] += lp;
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        a.w[d][P] = w[d];
        a.ldi[d][P] = ldi[d];
    }
}

template <class T, class A>
__global__ __launch_bounds__(256) void k_adj_rc_gather(ParamsT<Dual<A>> p, AdjRhieChow<T, A> a) {
    const int R = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (R >= p.N || !fluid_at(a.flag, R)) return;
    const int x = R % p.X, t = R / p.X, y = t % p.Y, z = t / p.Y;
    A linv[3] = {0, 0, 0}, lu[3] = {0, 0, 0}, lp = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int c = d >> 1, e = d ^ 1;
        linv[c] += (A)0.5 * a.ldi[d][R];
        const int P1 = step_of(p, x, y, z, d, 1);
        if (!fluid_at(a.flag, P1)) continue;
        // P1's face e reaches back to R: mean 1/D, (u_c(P1) + u_c(R)) / 2,
        // and p_R as that face's neighbour value
        linv[c] += (A)0.5 * a.ldi[e][P1];
        lu[c] += (A)0.5 * ((e & 1) == 0 ? (A)1 : (A)-1) * a.lam_bp[P1];
        const bool fF = fluid_at(a.flag, step_of(p, x, y, z, e, 1));
        const bool bF = fluid_at(a.flag, step_of(p, x, y, z, d, 2));
        double cQ, cP, cF, cB;
        rc_weights(fF, bF, cQ, cP, cF, cB);
        lp += (A)cQ * a.w[e][P1];
        // P1's face d (away from R) has R behind P1
        lp += (A)0.25 * a.w[d][P1];
        // the cell two out reaches R through its face e (neighbour P1)
        const int P2 = step_of(p, x, y, z, d, 2);
        if (P2 >= 0) lp -= (A)0.25 * a.w[e][P2];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a.lam_invD[c][R] += linv[c];
        a.lam_u[c][R] += lu[c];
    }
    a.lam_p[R] += lp;
}
