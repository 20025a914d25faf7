// Stretched collocated grid (GridType.CollNonUniform, FvmCollNonUniform.compute):
// device bodies included by uws_kernels.hip after the uniform-grid bodies,
// whose helpers (Cell, inreg, lin, bnd_at, les_term, ...) they use.
//
// The step is the uniform one (SIMPLE only); what changes:
//  - spacings come from the per-axis cell sizes f.h[a] (facePos[i+1] -
//    facePos[i]); a neighbour outside the grid takes the cell's own size;
//  - face values are linear interpolations with weights wP = h_Q / (h_P + h_Q)
//    for the cell and 1 - wP for the neighbour;
//  - momentum convection uses the Rhie-Chow face velocities that the previous
//    pressure precompute left in velFaceFluxField (f.flux, f.fluxB);
//  - the LES and wall-function kernels stay the uniform ones (the reference
//    binds the uniform LES shaders for this grid too, FvmSolverGpu.cs:522-552).
//
// The Jacobi sweeps keep the uniform layout and kernels:
//  - momentum: S holds ds_a F_f / (h_P + h_Q) on the owned + face, so that the
//    coefficient of u_Q in P's equation is -h_P S and that of u_P in Q's
//    equation +h_Q S (the two interpolation weights); the sweep multiplies by
//    the cell's own size per axis instead of ds/4;
//  - pressure: cp holds DInv_f / distance, symmetric per face; each cell's
//    factor dv_P / den is folded into its 1/Dp (times dv/den) and bp (divided
//    by dv/den), so the one- and two-sweep pressure kernels run unchanged.
//    Dp keeps the reference's value (residual norms); the field download of
//    bp unfolds it.
#pragma once

template <class T>
struct NuGeo {
    T d[3];   // dx dy dz of the cell
    T ds[3];  // dsx = dy dz, dsy = dx dz, dsz = dx dy
    T dv;     // dx dy dz
    T dn[6];  // neighbour sizes E W N S U D (the cell's own outside the grid)
};

// FvmCollNonUniform.compute:130-157
template <class T>
__device__ __forceinline__ NuGeo<T> nu_geo(const ParamsT<T>& p, const FieldsT<T>& f, int x, int y, int z) {
    NuGeo<T> g;
    const T* __restrict__ hx = f.h[0];
    const T* __restrict__ hy = f.h[1];
    const T* __restrict__ hz = f.h[2];
    g.d[0] = hx[x];
    g.d[1] = hy[y];
    g.d[2] = hz[z];
    g.ds[0] = g.d[1] * g.d[2];
    g.ds[1] = g.d[0] * g.d[2];
    g.ds[2] = g.d[0] * g.d[1];
    g.dv = g.d[0] * g.d[1] * g.d[2];
    g.dn[E_] = x + 1 < p.X ? hx[x + 1] : g.d[0];
    g.dn[W_] = x > 0 ? hx[x - 1] : g.d[0];
    g.dn[N_] = y + 1 < p.Y ? hy[y + 1] : g.d[1];
    g.dn[S_] = y > 0 ? hy[y - 1] : g.d[1];
    g.dn[U_] = z + 1 < p.Z ? hz[z + 1] : g.d[2];
    g.dn[D_] = z > 0 ? hz[z - 1] : g.d[2];
    return g;
}

// size of the cell two away in direction d (FvmCollNonUniform.compute:970-983)
template <class T>
__device__ __forceinline__ T nu_size2(const ParamsT<T>& p, const FieldsT<T>& f, const NuGeo<T>& g, int x, int y,
                                      int z, int d) {
    const int a = d >> 1;
    const int c = a == 0 ? x : (a == 1 ? y : z);
    const int n = a == 0 ? p.X : (a == 1 ? p.Y : p.Z);
    const int i = c + ((d & 1) ? -2 : 2);
    return (i >= 0 && i < n) ? f.h[a][i] : g.d[a];
}

// velFaceFluxField value of face d of cell P: + faces live at the
// neighbour (or in the boundary planes), - faces at the cell itself
template <class T>
__device__ __forceinline__ T nu_face_flux(const ParamsT<T>& p, const FieldsT<T>& f, int x, int y, int z, int P,
                                          int d) {
    switch (d) {
    case E_: return x + 1 < p.X ? f.flux[0][P + 1] : f.fluxB[0][z * p.Y + y];
    case W_: return f.flux[0][P];
    case N_: return y + 1 < p.Y ? f.flux[1][P + p.X] : f.fluxB[1][z * p.X + x];
    case S_: return f.flux[1][P];
    case U_: return z + 1 < p.Z ? f.flux[2][P + p.X * p.Y] : f.fluxB[2][y * p.X + x];
    default: return f.flux[2][P];
    }
}

// face viscosities (FvmCollNonUniform.compute:482-503)
template <class T>
__device__ __forceinline__ void nu_face_vis(const ParamsT<T>& p, const FieldsT<T>& f, const NuGeo<T>& g, int x,
                                            int y, int z, int P, T vis[6]) {
    if (!p.useLES) {
#pragma unroll
        for (int d = 0; d < 6; ++d) vis[d] = p.nu;
        return;
    }
    const T nutP = f.nut[P];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        vis[d] = p.nu + nutP;
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (fluid(f, Q)) {
                const T wP = g.dn[d] / (g.d[d >> 1] + g.dn[d]);
                const T wN = (T)1.0f - wP;
                vis[d] = p.nu + wP * nutP + wN * f.nut[Q];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// CSVelPredictPreCompute (FvmCollNonUniform.compute:111-721) -> D, 1/D, b,
// and the owned E/N/U face coefficients of the momentum sweep.
// ---------------------------------------------------------------------------
template <class T>
__device__ void body_nu_vel_precompute(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    T* const* un = f.vel[r.state];
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    const NuGeo<T> g = nu_geo(p, f, x, y, z);
    T vis[6];
    nu_face_vis(p, f, g, x, y, z, P, vis);
    {
        const int nbr[3] = {x + 1 < p.X ? P + 1 : -1, y + 1 < p.Y ? P + p.X : -1, z + 1 < p.Z ? P + p.X * p.Y : -1};
        const int dirs[3] = {E_, N_, U_};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T Sv = 0, Gv = 0;
            const int Q = nbr[a];
            if (isF && Q >= 0 && fluid(f, Q)) {
                const T hQ = g.dn[dirs[a]];
                Sv = g.ds[a] * f.flux[a][Q] / (g.d[a] + hQ);  // the + face of P is the - face of Q
                Gv = vis[dirs[a]] * g.ds[a] / ((T)0.5f * g.d[a] + (T)0.5f * hQ);
            }
            f.S[a][P] = Sv;
            f.G[a][P] = Gv;
        }
    }
    if (!isF) {
        f.invD[0][P] = 0; f.invD[1][P] = 0; f.invD[2][P] = 0;
        return;
    }
    T Dx = g.dv / p.dt, Dy = g.dv / p.dt, Dz = g.dv / p.dt;
    T bU = 0, bV = 0, bW = 0;
    const V3<T> uP = ld3(un, P);
    // advection with the stored face velocities (:185-325)
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int a = d >> 1;
        const bool plus = (d & 1) == 0;
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        const T F = nu_face_flux(p, f, x, y, z, P, d);
        const T dsa = g.ds[a];
        if (inreg(p, nx, ny, nz)) {
            if (fluid(f, lin(p, nx, ny, nz))) {
                const T wP = g.dn[d] / (g.d[a] + g.dn[d]);
                const T t = dsa * wP * F;
                if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
            }
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const V3<T> ub = bnd_at(p, f, d, x, y, z);
                if (plus) { bU -= dsa * ub.x * F; bV -= dsa * ub.y * F; bW -= dsa * ub.z * F; }
                else { bU += dsa * ub.x * F; bV += dsa * ub.y * F; bW += dsa * ub.z * F; }
            } else if (p.velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                const T t = dsa * F;
                if (plus) { Dx += t; Dy += t; Dz += t; } else { Dx -= t; Dy -= t; Dz -= t; }
            }
        }
    }
    // pressure face values (:327-462)
    const T pP = f.p[P];
    T pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (fluid(f, Q)) {
                const T wP = g.dn[d] / (g.d[d >> 1] + g.dn[d]);
                const T wN = (T)1.0f - wP;
                pf[d] = wP * pP + wN * f.p[Q];
            } else {
                pf[d] = pP;
            }
        } else {
            pf[d] = p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE ? (T)0 : pP;
        }
    }
    bU += -((T)1.0f / p.den) * g.ds[0] * (pf[E_] - pf[W_]);
    bV += -((T)1.0f / p.den) * g.ds[1] * (pf[N_] - pf[S_]);
    bW += -((T)1.0f / p.den) * g.ds[2] * (pf[U_] - pf[D_]);
    // diffusion (:464-697)
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int a = d >> 1;
        const T dsa = g.ds[a];
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const T t = fluid(f, lin(p, nx, ny, nz)) ? vis[d] * dsa / ((T)0.5f * g.d[a] + (T)0.5f * g.dn[d])
                                                     : vis[d] * dsa / ((T)0.5f * g.d[a]);
            Dx += t; Dy += t; Dz += t;
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const V3<T> ub = bnd_at(p, f, d, x, y, z);
                const T k = vis[d] * dsa / ((T)0.5f * g.d[a]);
                bU += k * ub.x; bV += k * ub.y; bW += k * ub.z;
                Dx += k; Dy += k; Dz += k;
            } else if (p.velBnd[bf] == UWS_VEL_SYMMETRY) {
                const T k = vis[d] * dsa / ((T)0.5f * g.d[a]);
                if (a == 0) Dx += k; else if (a == 1) Dy += k; else Dz += k;
            }
        }
    }
    // time derivative, force, LES (:699-715)
    bU += g.dv / p.dt * uP.x; bV += g.dv / p.dt * uP.y; bW += g.dv / p.dt * uP.z;
    bU += g.dv * p.fext[0]; bV += g.dv * p.fext[1]; bW += g.dv * p.fext[2];
    if (p.useLES) {
        V3<T> bl;
        if (r.flags & K_FUSE_LES) bl = les_term(p, f, un, c, nbrs_of(p, f, c));
        else bl = mk<T>(f.b[0][P], f.b[1][P], f.b[2][P]);
        bU += bl.x; bV += bl.y; bW += bl.z;
    }
    f.D[0][P] = Dx; f.D[1][P] = Dy; f.D[2][P] = Dz;
    f.invD[0][P] = (T)1.0f / Dx; f.invD[1][P] = (T)1.0f / Dy; f.invD[2][P] = (T)1.0f / Dz;
    f.b[0][P] = bU; f.b[1][P] = bV; f.b[2][P] = bW;
}

// ---------------------------------------------------------------------------
// CSPresCorrectRhieChowPreCompute (FvmCollNonUniform.compute:920-1344) ->
// Dp, folded 1/Dp and bp, the owned face coefficients, and (unless
// K_NO_FLUX) the face velocities of velFaceFluxField.
// ---------------------------------------------------------------------------
template <class T>
__device__ void body_nu_pres_precompute(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    T* const* u = f.vel[r.vout];
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    const bool isF = fluid(f, P);
    const NuGeo<T> g = nu_geo(p, f, x, y, z);
    {
        const int nbr[3] = {x + 1 < p.X ? P + 1 : -1, y + 1 < p.Y ? P + p.X : -1, z + 1 < p.Z ? P + p.X * p.Y : -1};
        const int dirs[3] = {E_, N_, U_};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T cv = 0;
            const int Q = nbr[a];
            if (isF && Q >= 0 && fluid(f, Q)) {
                const T hQ = g.dn[dirs[a]];
                const T wP = hQ / (g.d[a] + hQ);
                const T wN = (T)1.0f - wP;
                const T DInv = wP * f.invD[a][P] + wN * f.invD[a][Q];
                cv = DInv / ((T)0.5f * g.d[a] + (T)0.5f * hQ);
            }
            f.cp[a][P] = cv;
        }
    }
    if (!isF) {
        f.invDp[P] = 0;
        f.bp[P] = 0;
        return;
    }
    const T dvden = g.dv / p.den;
    const T half = (T)0.5f, h3 = (T)1.5f;
    T Dp = 0, bp = 0;
    T uf[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int a = d >> 1;
        const bool plus = (d & 1) == 0;
        const int ox = c_off[d][0], oy = c_off[d][1], oz = c_off[d][2];
        const int nx = x + ox, ny = y + oy, nz = z + oz;
        const T uPc = u[a][P];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (!fluid(f, Q)) continue;  // a solid neighbour leaves the face velocity at 0
            const T hP = g.d[a], hQ = g.dn[d], hB = g.dn[d ^ 1];
            const T wP = hQ / (hP + hQ);
            const T wN = (T)1.0f - wP;
            const T DInv = wP * f.invD[a][P] + wN * f.invD[a][Q];
            Dp -= dvden * DInv / (half * hP + half * hQ);
            T face = wP * uPc + wN * u[a][Q];
            // Rhie-Chow correction, one-sided averages next to solids and walls
            const int fx = x + 2 * ox, fy = y + 2 * oy, fz = z + 2 * oz;
            const int bx = x - ox, by = y - oy, bz = z - oz;
            const bool farF = inreg(p, fx, fy, fz) && fluid(f, lin(p, fx, fy, fz));
            const bool backF = inreg(p, bx, by, bz) && fluid(f, lin(p, bx, by, bz));
            const T hF = nu_size2(p, f, g, x, y, z, d);
            const T pP = f.p[P], pQ = f.p[Q];
            T gr, ga;
            if (plus) {
                gr = (pQ - pP) / (half * hP + half * hQ);
                if (farF && backF) {
                    const T pF = f.p[lin(p, fx, fy, fz)], pB = f.p[lin(p, bx, by, bz)];
                    ga = half * ((pF - pP) / (half * hF + hQ + half * hP) + (pQ - pB) / (half * hQ + hP + half * hB));
                } else if (farF) {
                    const T pF = f.p[lin(p, fx, fy, fz)];
                    ga = half * ((pF - pP) / (half * hF + hQ + half * hP) + (pQ - pP) / (half * hQ + h3 * hP));
                } else if (backF) {
                    const T pB = f.p[lin(p, bx, by, bz)];
                    ga = half * ((pQ - pP) / (h3 * hQ + half * hP) + (pQ - pB) / (half * hQ + hP + half * hB));
                } else {
                    ga = gr;
                }
            } else {
                gr = (pP - pQ) / (half * hP + half * hQ);
                if (farF && backF) {
                    const T pF = f.p[lin(p, fx, fy, fz)], pB = f.p[lin(p, bx, by, bz)];
                    ga = half * ((pB - pQ) / (half * hB + hP + half * hQ) + (pP - pF) / (half * hP + hQ + half * hF));
                } else if (farF) {
                    const T pF = f.p[lin(p, fx, fy, fz)];
                    ga = half * ((pP - pQ) / (h3 * hP + half * hQ) + (pP - pF) / (half * hP + hQ + half * hF));
                } else if (backF) {
                    const T pB = f.p[lin(p, bx, by, bz)];
                    ga = half * ((pB - pQ) / (half * hB + hP + half * hQ) + (pP - pQ) / (half * hP + h3 * hQ));
                } else {
                    ga = gr;
                }
            }
            face -= (g.dv * DInv) / p.den * (gr - ga);
            if (plus) bp += face; else bp -= face;
            uf[d] = face;
        } else {
            const int bf = bface_of(d);
            if (p.velBnd[bf] == UWS_VEL_FIXED_VALUE) {
                const T ubc = comp(bnd_at(p, f, d, x, y, z), a);
                uf[d] = ubc;
                if (plus) bp += ubc; else bp -= ubc;
            } else if (p.velBnd[bf] == UWS_VEL_ZERO_GRAD) {
                uf[d] = uPc;
                if (plus) bp += uPc; else bp -= uPc;
            }
            if (p.presBnd[bf] == UWS_PRES_FIXED_VALUE) {
                const T DInv = f.invD[a][P];
                Dp -= dvden * DInv / (half * g.d[a]);
            }
        }
    }
    if (!(r.flags & K_NO_FLUX)) {
        // velFaceFluxField: W / S / D faces at the cell, + faces on the domain boundary (:1335-1342)
        f.flux[0][P] = uf[W_];
        f.flux[1][P] = uf[S_];
        f.flux[2][P] = uf[D_];
        if (x + 1 == p.X) f.fluxB[0][z * p.Y + y] = uf[E_];
        if (y + 1 == p.Y) f.fluxB[1][z * p.X + x] = uf[N_];
        if (z + 1 == p.Z) f.fluxB[2][y * p.X + x] = uf[U_];
    }
    f.Dp[P] = Dp;
    f.invDp[P] = (T)1.0f / Dp * dvden;
    f.bp[P] = bp / dvden;
}

// CSApplyVelCorrectionDivergenceTheorem (FvmCollNonUniform.compute:1487-1732)
template <class T>
__device__ void body_nu_vel_correct(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok) return;
    const int x = c.x, y = c.y, z = c.z, P = c.P;
    if (!fluid(f, P)) return;
    const NuGeo<T> g = nu_geo(p, f, x, y, z);
    const T* __restrict__ pc = f.pc[r.pin];
    const T pcP = pc[P];
    // K_APPLY_P: CSApplyPresCorrection's p += p' in the same launch
    if (r.flags & K_APPLY_P) f.p[P] = f.p[P] + pcP;
    T pf[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        const int nx = x + c_off[d][0], ny = y + c_off[d][1], nz = z + c_off[d][2];
        if (inreg(p, nx, ny, nz)) {
            const int Q = lin(p, nx, ny, nz);
            if (fluid(f, Q)) {
                const T wP = g.dn[d] / (g.d[d >> 1] + g.dn[d]);
                const T wN = (T)1.0f - wP;
                pf[d] = wP * pcP + wN * pc[Q];
            } else {
                pf[d] = pcP;
            }
        } else {
            pf[d] = p.presBnd[bface_of(d)] == UWS_PRES_FIXED_VALUE ? (T)0 : pcP;
        }
    }
    const T dvden = g.dv / p.den;
    const T cx = -dvden * ((T)1.0f / f.D[0][P]) * (pf[E_] - pf[W_]) / g.d[0];
    const T cy = -dvden * ((T)1.0f / f.D[1][P]) * (pf[N_] - pf[S_]) / g.d[1];
    const T cz = -dvden * ((T)1.0f / f.D[2][P]) * (pf[U_] - pf[D_]) / g.d[2];
    f.ucorr[0][P] = cx; f.ucorr[1][P] = cy; f.ucorr[2][P] = cz;
    if (r.flags & K_NO_INPLACE) return;
    T* const* uo = f.vel[r.vout];
    uo[0][P] += cx; uo[1][P] += cy; uo[2][P] += cz;
}

// squared Rhie-Chow net outflow, bp unfolded
template <class T>
__device__ double body_nu_divergence(const ParamsT<T>& p, const FieldsT<T>& f, const Cell& c, const Roles& r) {
    if (!c.ok || !fluid(f, c.P)) return 0.0;
    const T dvden = f.h[0][c.x] * f.h[1][c.y] * f.h[2][c.z] / p.den;
    const double v = (double)(f.bp[c.P] * dvden);
    return v * v;
}
