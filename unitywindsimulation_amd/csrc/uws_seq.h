// Per-face sequential forms of five cell kernels (LES term, momentum
// precompute, Rhie-Chow precompute, PISO A_od u' and PISO precompute), used
// for the complex-step and forward-mode dual types.  The forms in
// uws_kernels.hip load all six neighbours' values before using any of them,
// which keeps the loads of a plain-float cell in flight together; for a
// two-double type the same live set exceeds the register file (the momentum
// precompute needed 253 VGPRs and 200 bytes of scratch per lane).  These
// compute the same values with each neighbour's loads under its branch.
// Included by uws_kernels.hip after the helpers they use.
#pragma once

// face viscosities with eddy viscosity (FvmColl.compute:1596-1635; LES.compute:643-679)
template <class T>
__device__ __forceinline__ void face_vis_seq(const ParamsT<T>& p, const FieldsT<T>& f, int x, int y, int z, int P,
                                         T vis[6]) {
    const T nu = p.nu;
    const T nutP = f.nut[P];
    const T base = nu + nutP;
#pragma unroll
    for (int d = 0; d < 6; ++d) vis[d] = base;
    // E: a solid neighbour contributes its own .x slot; W: this cell's .x slot (N/S, U/D alike)
    if (x + 1 < p.X) {
        const int Q = P + 1;
        if (fluid(f, Q)) vis[E_] = nu + (T)0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[E_] = nu + f.nuw[0][Q];
    }
    if (x - 1 >= 0) {
        const int Q = P - 1;
        if (fluid(f, Q)) vis[W_] = nu + (T)0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[W_] = nu + f.nuw[0][P];
    }
    if (y + 1 < p.Y) {
        const int Q = P + p.X;
        if (fluid(f, Q)) vis[N_] = nu + (T)0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[N_] = nu + f.nuw[1][Q];
    }
    if (y - 1 >= 0) {
        const int Q = P - p.X;
        if (fluid(f, Q)) vis[S_] = nu + (T)0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[S_] = nu + f.nuw[1][P];
    } else if (p.velBnd[UWS_Y0] == UWS_VEL_FIXED_VALUE) {
        vis[S_] = nu + f.nuw[1][P];
    }
    const int sz = p.X * p.Y;
    if (z + 1 < p.Z) {
        const int Q = P + sz;
        if (fluid(f, Q)) vis[U_] = nu + (T)0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[U_] = nu + f.nuw[2][Q];
    }
    if (z - 1 >= 0) {
        const int Q = P - sz;
        if (fluid(f, Q)) vis[D_] = nu + (T)0.5f * (nutP + f.nut[Q]);
        else if (solid(f, Q)) vis[D_] = nu + f.nuw[2][P];
    }
}


// LES transposed-stress term of one fluid cell (LES.compute:326-683).  The
// cell gradients of this cell and its fluid neighbours are those the
// eddy-viscosity kernel computed from the same u^n (cell_grad), read back.
template <class T>
__device__ __forceinline__ V3<T> les_term_seq(const ParamsT<T>& p, const FieldsT<T>& f, T* const un[3], int x, int y,
                                          int z, int P) {
    const V3<T> uP = ld3(un, P);
    const T hdx = (T)0.5f * p.dx;
    T g[3][3];
    load_grad(f, P, g);
    T fg[6][3];
#pragma unroll
    for (int d = 0; d < 6; ++d) { fg[d][0] = 0; fg[d][1] = 0; fg[d][2] = 0; }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        const T uPc = comp(uP, cc);
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (fluid(f, Q)) {
                T gn[3][3];
                load_grad(f, Q, gn);
#pragma unroll
                for (int j = 0; j < 3; ++j) fg[d][j] = (T)0.5f * (g[cc][j] + gn[cc][j]);
            } else {
                fg[d][cc] = plus ? -uPc / hdx : uPc / hdx;
            }
        } else {
            const int t = p.velBnd[bface_of(d)];
            if (t == UWS_VEL_FIXED_VALUE) {
                const T ubc = comp(bnd_at(p, f, d, x, y, z), cc);
                fg[d][cc] = plus ? (ubc - uPc) / hdx : (uPc - ubc) / hdx;
            } else if (t == UWS_VEL_ZERO_GRAD || t == UWS_VEL_SYMMETRY) {
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (j != cc) fg[d][j] = g[cc][j];
            }
        }
    }
    T vis[6];
    face_vis_seq(p, f, x, y, z, P, vis);
    T rr[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        rr[i] = p.ds * (vis[0] * fg[0][i] - vis[1] * fg[1][i] + vis[2] * fg[2][i] - vis[3] * fg[3][i] +
                        vis[4] * fg[4][i] - vis[5] * fg[5][i]);
    return mk<T>(rr[0], rr[1], rr[2]);
}

// CSCalLesDeferCorrectTerm (LES.compute:326-687) -> bField
template <class T>
__device__ void body_les_term_seq(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok || !fluid(f, c.P)) return;
    const V3<T> v = les_term_seq(p, f, f.vel[r.state], c.x, c.y, c.z, c.P);
    f.b[0][c.P] = v.x; f.b[1][c.P] = v.y; f.b[2][c.P] = v.z;
}


// ---------------------------------------------------------------------------
// CSVelPredictPreCompute (FvmColl.compute:1291-1859), optionally fused with
// the LES term (which it would otherwise read back from bField).  Besides D
// and b it stores the sweep's face coefficients S, G of the owned E/N/U
// faces (zero unless both cells are fluid) and 1/D.
// ---------------------------------------------------------------------------
template <class T>
__device__ void body_vel_precompute_seq(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    T* const* un = f.vel[r.state];
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    const V3<T> uP = ld3(un, P);
    T vis[6];
    if (p.useLES && isF) face_vis_seq(p, f, x, y, z, P, vis);
    else {
#pragma unroll
        for (int d = 0; d < 6; ++d) vis[d] = p.nu;
    }
    {
        const int nbr[3] = {x + 1 < p.X ? P + 1 : -1, y + 1 < p.Y ? P + p.X : -1, z + 1 < p.Z ? P + p.X * p.Y : -1};
        const int dirs[3] = {E_, N_, U_};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T Sv = 0, Gv = 0;
            const int Q = nbr[a];
            if (isF && Q >= 0 && fluid(f, Q)) {
                Sv = comp(uP, a) + un[a][Q];
                Gv = vis[dirs[a]] * p.ds / p.dx;
            }
            f.S[a][P] = Sv;
            f.G[a][P] = Gv;
        }
    }
    if (!isF) {
        f.invD[0][P] = 0; f.invD[1][P] = 0; f.invD[2][P] = 0;
        return;
    }
    T Dx = p.dv / p.dt, Dy = p.dv / p.dt, Dz = p.dv / p.dt;
    T bU = 0, bV = 0, bW = 0;
    const T ds = p.ds;
    // advection (:1314-1482)
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        const T uPn = comp(uP, cc);
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (fluid(f, Q)) {
                const T t = ds * (T)0.25f * (uPn + un[cc][Q]);
                if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
            }
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const V3<T> ub = bnd_at(p, f, d, x, y, z);
                const T ubn = comp(ub, cc);
                if (plus) { bU -= ds * ub.x * ubn; bV -= ds * ub.y * ubn; bW -= ds * ub.z * ubn; }
                else { bU += ds * ub.x * ubn; bV += ds * ub.y * ubn; bW += ds * ub.z * ubn; }
            } else if (p.velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                const T t = ds * uPn;
                if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
            }
        }
    }
    // pressure face values (:1484-1591)
    const T pP = f.p[P];
    T pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            pf[d] = fluid(f, Q) ? (T)0.5f * (pP + f.p[Q]) : pP;
        } else {
            pf[d] = p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE ? (T)0 : pP;
        }
    }
    bU += -((T)1.0f / p.den) * ds * (pf[E_] - pf[W_]);
    bV += -((T)1.0f / p.den) * ds * (pf[N_] - pf[S_]);
    bW += -((T)1.0f / p.den) * ds * (pf[U_] - pf[D_]);
    // diffusion (:1593-1836)
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            const T t = fluid(f, Q) ? vis[d] * ds / p.dx : (T)2.0f * vis[d] * ds / p.dx;
            Dx += t; Dy += t; Dz += t;
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const V3<T> ub = bnd_at(p, f, d, x, y, z);
                bU += (T)2.0f * vis[d] * ds / p.dx * ub.x;
                bV += (T)2.0f * vis[d] * ds / p.dx * ub.y;
                bW += (T)2.0f * vis[d] * ds / p.dx * ub.z;
                const T t = (T)2.0f * vis[d] * ds / p.dx;
                Dx += t; Dy += t; Dz += t;
            } else if (p.velBnd[bf] == UWS_VEL_SYMMETRY) {
                const T t = (T)2.0f * vis[d] * ds / p.dx;
                const int cc = d >> 1;
                if (cc == 0) Dx += t; else if (cc == 1) Dy += t; else Dz += t;
            }
        }
    }
    // time derivative, force, LES (:1838-1854)
    bU += p.dv / p.dt * uP.x; bV += p.dv / p.dt * uP.y; bW += p.dv / p.dt * uP.z;
    bU += p.dv * p.fext[0]; bV += p.dv * p.fext[1]; bW += p.dv * p.fext[2];
    if (p.useLES) {
        V3<T> bl;
        if (r.flags & K_FUSE_LES) bl = les_term_seq(p, f, un, x, y, z, P);
        else bl = mk<T>(f.b[0][P], f.b[1][P], f.b[2][P]);
        bU += bl.x; bV += bl.y; bW += bl.z;
    }
    f.D[0][P] = Dx; f.D[1][P] = Dy; f.D[2][P] = Dz;
    f.invD[0][P] = (T)1.0f / Dx; f.invD[1][P] = (T)1.0f / Dy; f.invD[2][P] = (T)1.0f / Dz;
    f.b[0][P] = bU; f.b[1][P] = bV; f.b[2][P] = bW;
}


// ---------------------------------------------------------------------------
// CSPresCorrectRhieChowPreCompute (FvmColl.compute:2480-2787) -> Dp, bp,
// 1/Dp, and the pressure face coefficients of the owned E/N/U faces.
// ---------------------------------------------------------------------------
template <class T>
__device__ void body_pres_precompute_seq(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    T* const* u = f.vel[r.vout];
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    const T dvden = p.dv / p.den;
    {
        const int nbr[3] = {x + 1 < p.X ? P + 1 : -1, y + 1 < p.Y ? P + p.X : -1, z + 1 < p.Z ? P + p.X * p.Y : -1};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T cv = 0, ce = 0;
            const int Q = nbr[a];
            if (isF && Q >= 0 && fluid(f, Q)) {
                const T DInv = (T)0.5f * (f.invD[a][P] + f.invD[a][Q]);
                cv = dvden * DInv / p.dx;
                ce = dvden * (DInv / p.dx);
            }
            f.cp[a][P] = cv;
            if (a == 0) f.cpisoE[P] = ce;
        }
    }
    if (!isF) {
        f.invDp[P] = 0;
        f.bp[P] = 0;
        return;
    }
    const T two_dx = (T)2.0f * p.dx;
    T Dp = 0, bp = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const bool plus = (d & 1) == 0;
        const int ox = c_off[d][0], oy = c_off[d][1], oz = c_off[d][2];
        const int nx = x + ox, ny = y + oy, nz = z + oz;
        const T uPc = u[cc][P];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (!fluid(f, Q)) continue;
            const T DInv = (T)0.5f * (f.invD[cc][P] + f.invD[cc][Q]);
            Dp -= dvden * DInv / p.dx;
            const T half = (T)0.5f * (uPc + u[cc][Q]);
            if (plus) bp += half; else bp -= half;
            const int fx = x + 2 * ox, fy = y + 2 * oy, fz = z + 2 * oz;
            const int bx = x - ox, by = y - oy, bz = z - oz;
            const bool farF = inreg(p, fx, fy, fz) && fluid(f, lin(p, fx, fy, fz));
            const bool backF = inreg(p, bx, by, bz) && fluid(f, lin(p, bx, by, bz));
            const T pP = f.p[P], pQ = f.p[Q];
            T g, ga;
            if (plus) {
                g = (pQ - pP) / p.dx;
                if (farF && backF) {
                    const T pF = f.p[lin(p, fx, fy, fz)], pB = f.p[lin(p, bx, by, bz)];
                    ga = (T)0.5f * ((pF - pP) / two_dx + (pQ - pB) / two_dx);
                } else if (farF) {
                    const T pF = f.p[lin(p, fx, fy, fz)];
                    ga = (T)0.5f * ((pF - pP) / two_dx + (pQ - pP) / two_dx);
                } else if (backF) {
                    const T pB = f.p[lin(p, bx, by, bz)];
                    ga = (T)0.5f * ((pQ - pP) / two_dx + (pQ - pB) / two_dx);
                } else {
                    ga = g;
                }
            } else {
                g = (pP - pQ) / p.dx;
                if (farF && backF) {
                    const T pF = f.p[lin(p, fx, fy, fz)], pB = f.p[lin(p, bx, by, bz)];
                    ga = (T)0.5f * ((pB - pQ) / two_dx + (pP - pF) / two_dx);
                } else if (farF) {
                    const T pF = f.p[lin(p, fx, fy, fz)];
                    ga = (T)0.5f * ((pP - pQ) / two_dx + (pP - pF) / two_dx);
                } else if (backF) {
                    const T pB = f.p[lin(p, bx, by, bz)];
                    ga = (T)0.5f * ((pB - pQ) / two_dx + (pP - pQ) / two_dx);
                } else {
                    ga = g;
                }
            }
            const T vct = -(p.dv * DInv) / p.den * (g - ga);
            if (plus) bp += vct; else bp -= vct;
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const T ubc = comp(bnd_at(p, f, d, x, y, z), cc);
                if (plus) bp += ubc; else bp -= ubc;
            } else if (p.velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                if (plus) bp += uPc; else bp -= uPc;
            }
            if (p.presBnd[bf] == UWS_PRES_FIXED_VALUE) {
                const T DInv = (T)1.0f / f.D[cc][P];
                Dp -= (T)2.0f * dvden * DInv / p.dx;
            }
        }
    }
    f.Dp[P] = Dp;
    f.invDp[P] = (T)1.0f / Dp;
    f.bp[P] = bp;
}


// CSPisoCalAodUCorrect (PisoColl.compute:70-178): laminar nu, no BC terms
template <class T>
__device__ void body_piso_aodu_seq(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    if (!fluid(f, P)) return;
    T* const* un = f.vel[r.state];
    const V3<T> uP = ld3(un, P);
    T A[3] = {0, 0, 0}, Df[3] = {0, 0, 0};
    const T q = p.ds * (T)0.25f;
    int Qs[6];
    bool ok[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = c.x + c_off[d][0], ny = c.y + c_off[d][1], nz = c.z + c_off[d][2];
        ok[d] = inreg(p, nx, ny, nz) && fluid(f, lin(p, nx, ny, nz));
        Qs[d] = ok[d] ? lin(p, nx, ny, nz) : P;
    }
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!ok[d]) continue;
        const int cc = d >> 1;
        const int Q = Qs[d];
        const T sum = comp(uP, cc) + un[cc][Q];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T t = q * f.ucorr[a][Q] * sum;
            if ((d & 1) == 0) A[a] += t; else A[a] -= t;
        }
    }
    const T g = p.nu * p.ds / p.dx;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!ok[d]) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) Df[a] -= g * f.ucorr[a][Qs[d]];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) f.aodu[a][P] = A[a] + Df[a];
}


// CSPisoPresCorrection's constant parts (PisoColl.compute:196-255): presCoeff
// and velTerm do not change over the sweeps and are computed once.
template <class T>
__device__ void body_piso_precompute_seq(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int P = c.P;
    if (!fluid(f, P)) {
        f.invDpiso[P] = 0;
        f.bpiso[P] = 0;
        return;
    }
    T presCoeff = 0, velTerm = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int cc = d >> 1;
        const int nx = c.x + c_off[d][0], ny = c.y + c_off[d][1], nz = c.z + c_off[d][2];
        if (!(inreg(p, nx, ny, nz) && fluid(f, lin(p, nx, ny, nz)))) continue;
        const int Q = lin(p, nx, ny, nz);
        const T AdInv = (T)0.5f * (f.invD[cc][P] + f.invD[cc][Q]);
        const T AodUCorrect = (T)0.5f * (f.aodu[cc][P] + f.aodu[cc][Q]);
        // this face's coefficient as stored on the cell owning it
        const int owner = (d & 1) == 0 ? P : Q;
        const T cf = cc == 0 ? f.cpisoE[owner] : f.cp[cc][owner];
        presCoeff -= cf;
        if ((d & 1) == 0) velTerm -= AdInv * AodUCorrect; else velTerm += AdInv * AodUCorrect;
    }
    f.invDpiso[P] = (T)1.0f / presCoeff;
    f.bpiso[P] = velTerm;
}
