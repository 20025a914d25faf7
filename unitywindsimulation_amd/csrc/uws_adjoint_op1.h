// Reverse of the momentum coefficients (wall function, eddy viscosity, LES
// term and CSVelPredictPreCompute: u^n, p^n, boundary textures, Cs -> S, G,
// D, 1/D, b) by hand, in gather form.  Included by uws_kernels.hip inside its
// anonymous namespace, after the forward bodies whose helpers it uses.
//
// The primal values are those of one forward run of the same bodies in the
// adjoint arithmetic (the Dual mirror with zero tangents: its .v parts).
// Three kernels, one thread per cell, each writing only its own cell:
//  k_op1_local  for a fluid cell P, the adjoints of the quantities its own
//               outputs depend on: its six face viscosities, its six faces'
//               LES face gradients, the face advection and face pressure
//               weights, and what falls on its own u, p, cell gradient,
//               eddy viscosity and boundary-texture elements directly;
//  k_op1_grad   for a fluid cell R, the total adjoint of its eddy viscosity
//               (own faces plus the neighbours' faces towards R), of its
//               cell gradient (likewise, plus the eddy viscosity's
//               dependence on it) and its share of dJ/dCs;
//  k_op1_state  for a fluid cell R, the adjoint of u^n_R and p^n_R: its own
//               share plus every neighbour's face terms that read R, and the
//               face velocities of its own and its neighbours' gradients.
// Solid cells' u and p are read by no momentum coefficient: their adjoints
// get nothing here.
#pragma once

// dir 0..5 = E W N S U D; the face of a neighbour that faces back
__device__ __forceinline__ int opp_dir(int d) { return d ^ 1; }

// boundary-texture element of the face a boundary cell touches in direction d
template <class T>
__device__ __forceinline__ int bnd_elem(const ParamsT<T>& p, int d, int x, int y, int z) {
    const int bf = bface_of(d);
    if (bf < 2) return z * p.Y + y;
    if (bf < 4) return z * p.X + x;
    return y * p.X + x;
}

// d(nu_w)/d(u) of a wall face normal to axis a of fluid cell P with velocity u:
// the wall function of the speed along the face, differentiated in one
// forward-mode evaluation; g[c] gets the derivative w.r.t. u_c
template <class A>
__device__ __forceinline__ void wall_grad(const ParamsT<Dual<A>>& p, const A u[3], int a, A g[3]) {
    const int t0 = a == 0 ? 1 : 0, t1 = a == 2 ? 1 : 2;
    g[0] = g[1] = g[2] = 0;
    const A s = sqrt(u[t0] * u[t0] + u[t1] * u[t1]);
    if (!(s > (A)0)) return;
    const Dual<A> nw = log_law_wall_func<Dual<A>, false>(p, Dual<A>(s, (A)1));
    g[t0] = nw.d * u[t0] / s;
    g[t1] = nw.d * u[t1] / s;
}


template <class A>
__device__ __forceinline__ void face_vis_v(const ParamsT<Dual<A>>& p, const FieldsT<Dual<A>>& f, int x, int y, int z,
                                           int P, A vis[6]) {
    const A nu = p.nu.v;
    const A nutP = f.nut[P].v;
#pragma unroll
    for (int d = 0; d < 6; ++d) vis[d] = nu + nutP;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        const int a = d >> 1;
        const bool plus = (d & 1) == 0;
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (fluid(f, Q)) vis[d] = nu + (A)0.5 * (nutP + f.nut[Q].v);
            else if (solid(f, Q)) vis[d] = nu + f.nuw[a][plus ? Q : P].v;
        } else if (d == S_ && p.velBnd[UWS_Y0] == UWS_VEL_FIXED_VALUE) {
            vis[d] = nu + f.nuw[1][P].v;
        }
    }
}

template <class A>
__global__ __launch_bounds__(256) void k_op1_local(ParamsT<Dual<A>> p, FieldsT<Dual<A>> f, Op1Args<A> a) {
    const int P = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (P >= a.N) return;
    const int x = P % p.X, t = P / p.X, y = t % p.Y, z = t / p.Y;
    A lv[6], lf[6][3], la[6], lq[6], luP[3], lpP = 0, lgP[3][3], lnutP = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        lv[d] = la[d] = lq[d] = 0;
        lf[d][0] = lf[d][1] = lf[d][2] = 0;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        luP[i] = 0;
        lgP[i][0] = lgP[i][1] = lgP[i][2] = 0;
    }
    if (fluid(f, P)) {
        const A ds = p.ds.v, dx = p.dx.v, hdx = (A)0.5 * dx, den = p.den.v;
        A uP[3], lDt[3], lb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            uP[c] = f.vel[0][c][P].v;
            const A Dc = f.D[c][P].v;
            lDt[c] = a.lamD[c][P] - a.lamInvD[c][P] / (Dc * Dc);
            lb[c] = a.lamB[c][P];
            luP[c] = p.dv.v / p.dt.v * lb[c];  // time term dv/dt u^n
        }
        const A sumD = lDt[0] + lDt[1] + lDt[2];
        A vis[6];
        if (a.les) face_vis_v(p, f, x, y, z, P, vis);
        else {
#pragma unroll
            for (int d = 0; d < 6; ++d) vis[d] = p.nu.v;
        }
        A g[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) g[i][j] = f.grad[3 * i + j][P].v;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int cc = d >> 1;
            const bool plus = (d & 1) == 0;
            const A sgn = plus ? (A)1 : (A)-1;
            const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
            // the LES face gradient of the face-normal component (LES.compute:326-683)
            A fg[3] = {0, 0, 0};
            if (inreg(p, nx, ny, nz)) {
                const int Q = lin(p, nx, ny, nz);
                if (fluid(f, Q)) {
                    // advection D_c += sgn ds/4 (u_P,n + u_Q,n); diffusion D_c += vis ds/dx
                    const A m = sgn * ds * (A)0.25 * sumD;
                    luP[cc] += m;
                    la[d] = m;
                    lv[d] += ds / dx * sumD;
                    // face pressure (p_P + p_Q)/2 in b_n -= (ds/den) (pf+ - pf-)
                    lq[d] = -sgn * ds / den * lb[cc];
                    lpP += (A)0.5 * lq[d];
                    if (plus) {
                        // owned face coefficients S = u_P,n + u_Q,n, G = vis ds/dx
                        luP[cc] += a.lamS[cc][P];
                        lv[d] += a.lamG[cc][P] * ds / dx;
                    }
                    if (a.les)
#pragma unroll
                        for (int j = 0; j < 3; ++j) fg[j] = (A)0.5 * (g[cc][j] + f.grad[3 * cc + j][Q].v);
                } else {
                    lv[d] += (A)2 * ds / dx * sumD;       // wall half a cell away
                    lpP += -sgn * ds / den * lb[cc];      // face pressure p_P
                    fg[cc] = plus ? -uP[cc] / hdx : uP[cc] / hdx;
                }
            } else {
                const int bf = bface_of(d);
                const int tv = p.velBnd[bf];
                const int e = bnd_elem(p, d, x, y, z);
                if (tv == UWS_VEL_FIXED_VALUE) {
                    A ub[3], lub[3] = {0, 0, 0};
#pragma unroll
                    for (int c = 0; c < 3; ++c) ub[c] = f.bnd[bf][c][e].v;
                    // b_c -= sgn ds ub_n ub_c;  b_c += 2 vis ds/dx ub_c;  D_c += 2 vis ds/dx
                    A sb = 0;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        sb += lb[c] * ub[c];
                        lub[c] += -sgn * ds * ub[cc] * lb[c] + (A)2 * vis[d] * ds / dx * lb[c];
                    }
                    lub[cc] += -sgn * ds * sb;
                    lv[d] += (A)2 * ds / dx * (sb + sumD);
                    if (a.les) {
                        fg[cc] = plus ? (ub[cc] - uP[cc]) / hdx : (uP[cc] - ub[cc]) / hdx;
                    }
                    // the LES term's fixed-value face (its adjoint is added below, once lf is known)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        if (a.lamBnd[bf][c]) a.lamBnd[bf][c][e] += lub[c];
                } else if (tv == UWS_VEL_ZERO_GRAD) {
                    luP[cc] += sgn * ds * sumD;  // D_c += sgn ds u_P,n
                    if (a.les)
#pragma unroll
                        for (int j = 0; j < 3; ++j) fg[j] = j != cc ? g[cc][j] : (A)0;
                } else if (tv == UWS_VEL_SYMMETRY) {
                    lv[d] += (A)2 * ds / dx * lDt[cc];
                    if (a.les)
#pragma unroll
                        for (int j = 0; j < 3; ++j) fg[j] = j != cc ? g[cc][j] : (A)0;
                }
                if (p.presBnd[bf] != UWS_PRES_FIXED_VALUE) lpP += -sgn * ds / den * lb[cc];
            }
            if (a.les) {
                // b_i += ds sgn vis fg_i
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    lv[d] += ds * sgn * lb[i] * fg[i];
                    lf[d][i] = ds * sgn * vis[d] * lb[i];
                }
            }
        }
        if (a.les) {
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                const int cc = d >> 1;
                const bool plus = (d & 1) == 0;
                const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
                if (inreg(p, nx, ny, nz)) {
                    const int Q = lin(p, nx, ny, nz);
                    if (fluid(f, Q)) {
                        lnutP += (A)0.5 * lv[d];  // nu + (nut_P + nut_Q)/2; Q's half is its own gather
#pragma unroll
                        for (int j = 0; j < 3; ++j) lgP[cc][j] += (A)0.5 * lf[d][j];
                    } else {
                        if (solid(f, Q) && (a.slots == 3 || cc == 0)) {
                            A w[3];
                            wall_grad(p, uP, cc, w);  // nu + nu_w of this wall face
#pragma unroll
                            for (int c = 0; c < 3; ++c) luP[c] += lv[d] * w[c];
                        } else if (!solid(f, Q)) {
                            lnutP += lv[d];
                        }
                        luP[cc] += (plus ? (A)-1 : (A)1) / ((A)0.5 * p.dx.v) * lf[d][cc];
                    }
                } else {
                    const int bf = bface_of(d);
                    const int tv = p.velBnd[bf];
                    if (d == S_ && tv == UWS_VEL_FIXED_VALUE) {
                        if (a.slots == 3) {
                            A w[3];
                            wall_grad(p, uP, 1, w);  // the ground's wall value
#pragma unroll
                            for (int c = 0; c < 3; ++c) luP[c] += lv[d] * w[c];
                        }
                    } else {
                        lnutP += lv[d];
                    }
                    if (tv == UWS_VEL_FIXED_VALUE) {
                        const int e = bnd_elem(p, d, x, y, z);
                        const A s1 = (plus ? (A)1 : (A)-1) / ((A)0.5 * p.dx.v) * lf[d][cc];
                        if (a.lamBnd[bf][cc]) a.lamBnd[bf][cc][e] += s1;
                        luP[cc] -= s1;
                    } else if (tv == UWS_VEL_ZERO_GRAD || tv == UWS_VEL_SYMMETRY) {
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            if (j != cc) lgP[cc][j] += lf[d][j];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        a.lvis[d][P] = lv[d];
        a.ladv[d][P] = la[d];
        a.lpf[d][P] = lq[d];
#pragma unroll
        for (int j = 0; j < 3; ++j) a.lfg[3 * d + j][P] = lf[d][j];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) a.lu[c][P] = luP[c];
    a.lp[P] = lpP;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) a.lg[3 * i + j][P] = lgP[i][j];
    a.lnut[P] = lnutP;
}

// eddy viscosity and cell gradient adjoints of fluid cell R (LES only)
template <class A>
__global__ __launch_bounds__(256) void k_op1_grad(ParamsT<Dual<A>> p, FieldsT<Dual<A>> f, Op1Args<A> a) {
    const int R = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (R >= a.N) return;
    A lcs = 0;
    if (fluid(f, R)) {
        const int x = R % p.X, t = R / p.X, y = t % p.Y, z = t / p.Y;
        A lnut = a.lnut[R];
        A lg[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) lg[i][j] = a.lg[3 * i + j][R];
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
            if (!inreg(p, nx, ny, nz)) continue;
            const int Q = lin(p, nx, ny, nz);
            if (!fluid(f, Q)) continue;
            const int b = opp_dir(d), cc = d >> 1;
            lnut += (A)0.5 * a.lvis[b][Q];
#pragma unroll
            for (int j = 0; j < 3; ++j) lg[cc][j] += (A)0.5 * a.lfg[3 * b + j][Q];
        }
        // nu_t = (Cs dx)^2 |S|, |S| = sqrt(2 S:S), d|S|/dg_ij = 2 S_ij / |S|
        A g[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) g[i][j] = f.grad[3 * i + j][R].v;
        A S[3][3], ss = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                S[i][j] = (A)0.5 * (g[i][j] + g[j][i]);
                ss += S[i][j] * S[i][j];
            }
        const A Sabs = sqrt((A)2 * ss);
        const A cs = p.cs.v, dx = p.dx.v;
        if (Sabs > (A)0) {
            const A k = lnut * (cs * cs) * (dx * dx) * (A)2 / Sabs;
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) lg[i][j] += k * S[i][j];
        }
        lcs = lnut * (A)2 * cs * (dx * dx) * Sabs;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) a.lg[3 * i + j][R] = lg[i][j];
    }
    a.lcs[R] = lcs;
}

template <class A>
__global__ __launch_bounds__(256) void k_op1_state(ParamsT<Dual<A>> p, FieldsT<Dual<A>> f, Op1Args<A> a) {
    const int R = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (R >= a.N || !fluid(f, R)) return;
    const int x = R % p.X, t = R / p.X, y = t % p.Y, z = t / p.Y;
    const A dx = p.dx.v;
    A lu[3] = {a.lu[0][R], a.lu[1][R], a.lu[2][R]};
    A lp = a.lp[R];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (!inreg(p, nx, ny, nz)) continue;
        const int Q = lin(p, nx, ny, nz);
        if (!fluid(f, Q)) continue;
        const int b = opp_dir(d);
        lu[cc] += a.ladv[b][Q];            // Q's advection of its face towards R
        lp += (A)0.5 * a.lpf[b][Q];       // Q's face pressure (p_Q + p_R)/2
        if (!plus) lu[cc] += a.lamS[cc][Q];  // S of Q's owned face = u_Q,n + u_R,n
        if (a.les) {
            // Q's gradient: g_Q[i][cc] = (fv(Q, +cc) - fv(Q, -cc))/dx, the face towards R
            // averaging u_Q and u_R; it is Q's + face when R lies on Q's + side
            const A s = plus ? (A)-1 : (A)1;  // R = Q - e (plus) reads Q's - face
#pragma unroll
            for (int i = 0; i < 3; ++i) lu[i] += s * (A)0.5 * a.lg[3 * i + cc][Q] / dx;
        }
    }
    if (a.les) {
        // R's own gradient: the face velocities of its six faces
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const int cc = d >> 1;
            const bool plus = (d & 1) == 0;
            const A s = plus ? (A)1 : (A)-1;
            const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
            if (inreg(p, nx, ny, nz)) {
                const int Q = lin(p, nx, ny, nz);
                if (fluid(f, Q))
#pragma unroll
                    for (int i = 0; i < 3; ++i) lu[i] += s * (A)0.5 * a.lg[3 * i + cc][R] / dx;
                // a wall face has the velocity 0
            } else {
                const int bf = bface_of(d);
                const int tv = p.velBnd[bf];
                if (tv == UWS_VEL_FIXED_VALUE) {
                    const int e = bnd_elem(p, d, x, y, z);
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        if (a.lamBnd[bf][i]) a.lamBnd[bf][i][e] += s * a.lg[3 * i + cc][R] / dx;
                } else {
#pragma unroll
                    for (int i = 0; i < 3; ++i)
                        if (tv == UWS_VEL_ZERO_GRAD || i != cc) lu[i] += s * a.lg[3 * i + cc][R] / dx;
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) a.lamU[c][R] += lu[c];
    a.lamP[R] += lp;
}
