// Discrete adjoint of nsteps solver steps with checkpointing
// (uws_adjoint_* entry points).
//
// Forward: the handle steps as usual (captured graphs) and (u, p) is copied
// every K steps.  Backward, segment by segment from the last: the segment is
// recomputed from its checkpoint by a recording step that runs the same
// kernels as uws_step on the same inputs (so the same values) but keeps
// every intermediate field and every Jacobi iterate in a per-step record;
// then its steps are reversed, last first.  One step of the forward is
//
//   Op1   (u^n, p^n, bnd, Cs) -> S, G, D, 1/D, b      wall fn + eddy vis + LES + momentum coefficients
//   V     u^{k+1} = invD (A(S, G) u^k + b)           momentum Jacobi sweeps, u* = u^V
//   Op3   (u*, p^n, 1/D, D, bnd) -> cp, cpE, 1/Dp, bp Rhie-Chow coefficients
//   P     p'^{k+1} = invDp (-C p'^k + bp)            SIMPLE pressure sweeps, p'^0 = 0
//   corr  p = p^n + p', u = u* + u'(p', D)
//   C x   aodu(u^n, u'), PISO coefficients, P sweeps, p += p', u += u'(p', 1/D, aodu)
//
// and its reverse walks this list backwards.  The sweeps are reversed by the
// transposed-sweep kernels, every one-shot kernel by a hand-written gather
// kernel (uws_adjoint_op1.h, uws_adjoint_ops.h) or, with
// UWS_ADJOINT_REVERSE=colour, by colouring + Dual runs of its own forward body
// (uws_adjoint.h).  In-place updates (u += u', p += p') are identities
// handled here directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <new>
#include <type_traits>
#include <vector>

#include "uws_adjoint.h"
#include "uws_host.h"
#include "uws_solver.h"

namespace uws {

struct AdjointBase {
    virtual ~AdjointBase() {}
    virtual int forward() = 0;
    virtual int backward(const double* lam_u, int64_t n_u, const double* lam_p, int64_t n_p) = 0;
    virtual int get_grad(int what, int face, double* dst, int64_t n) = 0;
    int device = 0;
    size_t bytes = 0;
};

template <class T>
ParamsT<Dual<T>> dual_params(const ParamsT<T>& p) {
    ParamsT<Dual<T>> d{};
    d.X = p.X; d.Y = p.Y; d.Z = p.Z; d.N = p.N; d.ty = p.ty; d.tiles = p.tiles;
    d.dx = p.dx; d.dt = p.dt; d.ds = p.ds; d.dv = p.dv; d.nu = p.nu; d.den = p.den; d.cs = p.cs;
    for (int i = 0; i < 3; ++i) d.fext[i] = p.fext[i];
    for (int i = 0; i < 6; ++i) { d.velBnd[i] = p.velBnd[i]; d.presBnd[i] = p.presBnd[i]; }
    d.useLES = p.useLES; d.wallSlots = p.wallSlots; d.wallModel = p.wallModel; d.wallZ0 = p.wallZ0;
    d.normRef = p.normRef;
    d.zchunk = p.zchunk;
    return d;
}

template <class T>
struct Adjoint : AdjointBase {
    using A = double;  // adjoint arithmetic (the records stay in T)
    using V3p = std::array<T*, 3>;
    using V3a = std::array<A*, 3>;
    Solver<T>* s = nullptr;
    int nsteps = 0, K = 1, nseg = 0;
    int V = 0, P = 0, C = 0;
    bool les = false;
    int X = 0, Y = 0, Z = 0, N = 0;
    hipStream_t st = nullptr;
    void* slab = nullptr;
    void* dslab = nullptr;
    bool fwd_done = false, have_grad = false;

    // one step's record
    struct Tape {
        std::vector<V3p> U;               // momentum iterates u^0 = u^n .. u^V
        T* P0 = nullptr;                  // p^n
        V3p S{}, G{}, D{}, invD{}, b{};   // Op1 outputs
        V3p cp{};
        T *cpe = nullptr, *Dp = nullptr, *invDp = nullptr, *bp = nullptr;
        std::vector<std::vector<T*>> PS;  // [solve][sweep] pressure iterates
        std::vector<V3p> UC;              // u' written by correction c (0 = SIMPLE)
        std::vector<V3p> AO;              // aodu of corrector c (1..C)
        std::vector<T*> IDPI, BPI;        // PISO 1/diag and right-hand side of corrector c
    };
    std::vector<Tape> tapes;
    std::vector<std::array<T*, 4>> ckpt;  // per segment: u (3), p
    T* zero = nullptr;
    T* w_nut = nullptr;
    V3p w_nuw{};
    V3p seg_u{};
    T* seg_p = nullptr;
    // adjoint fields
    V3a lamU[2]{};
    int cur = 0;
    A *lamP = nullptr, *lamPC[2] = {nullptr, nullptr};
    V3a lamInvD{}, lamD{}, lamS{}, lamG{}, lamB{}, lamCP{}, lamU0{}, lamUc{}, lamUcNew{}, lamUT{}, lamAO{};
    A *lamCPE = nullptr, *lamInvDp = nullptr, *lamBp = nullptr, *lamIDPI = nullptr, *lamBPI = nullptr;
    A* tcon = nullptr;
    // adjoints of the iterates of a window of transposed sweeps
    A* pwin[ADJ_WIN + 1] = {};
    V3a vwin[ADJ_WIN + 1]{};
    // the one-shot kernels are reversed by hand (adj_op1, adj_rhie_chow,
    // adj_correct, adj_piso_coef, adj_aodu) unless UWS_ADJOINT_REVERSE=colour
    // selects the colouring and Dual runs of their forward bodies (the two
    // are compared in the tests)
    bool hand = true;
    A* op1s[51] = {};  // scratch of the hand-written reverses
    double* sum_part = nullptr;
    V3a lamBnd[6]{};
    double* lamCs = nullptr;
    // forward-mode mirror of the fields, in the adjoint arithmetic
    ParamsT<Dual<A>> pd{};
    FieldsT<Dual<A>> fd{};

    ~Adjoint() override {
        if (st) (void)hipStreamSynchronize(st);
        if (slab) (void)hipFree(slab);
        if (dslab) (void)hipFree(dslab);
    }

    int bw(int face) const { return bnd_w(X, Y, face); }
    int bh(int face) const { return bnd_h(Y, Z, face); }

    // assign every device array from `base` (nullptr: only count); returns bytes
    size_t layout(char* base) {
        size_t off = 0;
        auto take = [&](size_t bytes) -> char* {
            off = (off + 255) & ~(size_t)255;
            char* r = base ? base + off : nullptr;
            off += bytes;
            return r;
        };
        auto pl = [&]() { return (T*)take((size_t)N * sizeof(T)); };
        auto v3 = [&]() { return V3p{pl(), pl(), pl()}; };
        auto pa = [&]() { return (A*)take((size_t)N * sizeof(A)); };
        auto a3 = [&]() { return V3a{pa(), pa(), pa()}; };
        tapes.assign(K, Tape{});
        for (auto& tp : tapes) {
            tp.U.resize(V + 1);
            for (auto& u : tp.U) u = v3();
            tp.P0 = pl();
            tp.S = v3(); tp.G = v3(); tp.D = v3(); tp.invD = v3(); tp.b = v3();
            tp.cp = v3();
            tp.cpe = pl(); tp.Dp = pl(); tp.invDp = pl(); tp.bp = pl();
            tp.PS.assign(C + 1, std::vector<T*>(P));
            for (auto& ps : tp.PS)
                for (auto& x : ps) x = pl();
            tp.UC.resize(C + 1);
            for (auto& u : tp.UC) u = v3();
            tp.AO.resize(C + 1);
            tp.IDPI.assign(C + 1, nullptr);
            tp.BPI.assign(C + 1, nullptr);
            for (int c = 1; c <= C; ++c) {
                tp.AO[c] = v3();
                tp.IDPI[c] = pl();
                tp.BPI[c] = pl();
            }
        }
        ckpt.resize(nseg);
        for (auto& c : ckpt) c = {pl(), pl(), pl(), pl()};
        zero = pl();
        w_nut = pl();
        w_nuw = v3();
        seg_u = v3();
        seg_p = pl();
        lamU[0] = a3(); lamU[1] = a3();
        lamP = pa(); lamPC[0] = pa(); lamPC[1] = pa();
        lamInvD = a3(); lamD = a3(); lamS = a3(); lamG = a3(); lamB = a3(); lamCP = a3();
        lamU0 = a3(); lamUc = a3(); lamUcNew = a3(); lamUT = a3(); lamAO = a3();
        lamCPE = pa(); lamInvDp = pa(); lamBp = pa(); lamIDPI = pa(); lamBPI = pa();
        tcon = pa();
        for (auto& w : pwin) w = pa();
        for (auto& w : vwin) w = a3();
        for (auto& q : op1s) q = pa();
        for (int f = 0; f < 6; ++f)
            for (int c = 0; c < 3; ++c) lamBnd[f][c] = (A*)take((size_t)bw(f) * bh(f) * sizeof(A));
        lamCs = (double*)take(sizeof(double));
        sum_part = (double*)take(ADJ_SUM_PARTS * sizeof(double));
        return off + 256;
    }

    int init(Solver<T>* solver, int n, int k) {
        s = solver;
        nsteps = n;
        K = std::min(k, n);
        nseg = (n + K - 1) / K;
        V = s->sc.V; P = s->sc.P; C = s->sc.C; les = s->sc.les;
        if (const char* e = std::getenv("UWS_ADJOINT_REVERSE")) hand = std::strcmp(e, "colour") != 0;
        X = s->p.X; Y = s->p.Y; Z = s->p.Z; N = s->p.N;
        st = s->stream;
        device = s->device;
        const size_t nb = layout(nullptr);
        if (hipMalloc(&slab, nb) != hipSuccess) { slab = nullptr; return UWS_ERR_ALLOC; }
        layout((char*)slab);
        set_dual_params();
        const size_t db = carve(pd, fd, nullptr, false);
        if (hipMalloc(&dslab, db) != hipSuccess) { dslab = nullptr; return UWS_ERR_ALLOC; }
        carve(pd, fd, (char*)dslab, false);
        fd.flag = s->f.flag;
        bytes = nb + db;
        if (hipMemsetAsync(slab, 0, nb, st) != hipSuccess || hipMemsetAsync(dslab, 0, db, st) != hipSuccess)
            return UWS_ERR_DEVICE;
        return s->sync();
    }

    // the solver's parameters in the adjoint arithmetic (Cs may have been
    // changed on the handle since creation)
    void set_dual_params() {
        ParamsT<A> pa{};
        fill_params(s->cfg, pa);
        pa.cs = (A)s->p.cs;
        pd = dual_params(pa);
    }

    // ---- helpers ----
    template <class E>
    int copy(E* dst, const E* src, size_t n = 0) {
        return hipMemcpyAsync(dst, src, (n ? n : (size_t)N) * sizeof(E), hipMemcpyDeviceToDevice, st) == hipSuccess
                   ? UWS_OK
                   : UWS_ERR_DEVICE;
    }
    template <class E>
    int clear(E* dst, size_t n = 0) {
        return hipMemsetAsync(dst, 0, (n ? n : (size_t)N) * sizeof(E), st) == hipSuccess ? UWS_OK : UWS_ERR_DEVICE;
    }
    // zero cell-sized adjoint arrays in one launch
    void zero_all(std::initializer_list<A*> arrays) {
        ZeroList<A> z{};
        for (A* a : arrays) {
            if (z.m == ADJ_ZERO_MAX) {
                adj_zero(z, N, st);
                z.m = 0;
            }
            z.p[z.m++] = a;
        }
        adj_zero(z, N, st);
    }
    void run(Kernel k, const FieldsT<T>& f, const Roles& r) { launch<T>(k, s->p, f, r, st); }

    // ---- forward with checkpoints ----
    int forward() override {
        if (s->rs.cur >= 0) return UWS_ERR_INVALID_ARG;  // a staged step is in progress
        for (int n = 0; n < nsteps; ++n) {
            if (n % K == 0) {
                auto& c = ckpt[n / K];
                for (int a = 0; a < 3; ++a)
                    if (int rc = copy(c[a], s->f.vel[s->rs.state][a])) return rc;
                if (int rc = copy(c[3], s->f.p)) return rc;
            }
            if (int rc = s->step_n(1)) return rc;
        }
        if (int rc = s->sync()) return rc;
        fwd_done = true;
        have_grad = false;
        return UWS_OK;
    }

    T* pc_final(const Tape& tp, int solve) const { return P > 0 ? tp.PS[solve][P - 1] : zero; }

    // one pressure solve with every iterate kept
    // (FP32 with the step's three-sweep kernels: three sweeps per launch,
    // the kernel storing its first two sweeps' iterates as well; the values
    // equal single sweeps')
    void record_pres(FieldsT<T> f, const Tape& tp, int solve) {
        for (int k = 0; k < P;) {
            const bool three = std::is_same_v<T, float> && use_fusion3(s->fuse_mode) && k + 3 <= P;
            f.pc[0] = k > 0 ? tp.PS[solve][k - 1] : zero;
            f.pc[1] = tp.PS[solve][k + (three ? 2 : 0)];
            f.pcrec[0] = three ? tp.PS[solve][k] : nullptr;
            f.pcrec[1] = three ? tp.PS[solve][k + 1] : nullptr;
            Roles r{};
            r.coef = solve > 0 ? 1 : 0;
            r.pin = k > 0 ? 0 : -1;
            r.pout = 1;
            run(three ? KN_PRES_SWEEP3 : KN_PRES_SWEEP, f, r);
            k += three ? 3 : 1;
        }
    }

    // the step of enqueue_step() from tp.U[0], tp.P0 into (out_u, out_p)
    int record_step(const Tape& tp, const V3p& out_u, T* out_p) {
        FieldsT<T> f = s->f;  // flags, boundary textures, scratch scalars
        for (int a = 0; a < 3; ++a) {
            f.vel[0][a] = tp.U[0][a];
            f.S[a] = tp.S[a]; f.G[a] = tp.G[a]; f.D[a] = tp.D[a]; f.invD[a] = tp.invD[a]; f.b[a] = tp.b[a];
            f.nuw[a] = w_nuw[a];
            f.cp[a] = tp.cp[a];
        }
        f.p = tp.P0;
        f.nut = w_nut;
        f.cpisoE = tp.cpe; f.Dp = tp.Dp; f.invDp = tp.invDp; f.bp = tp.bp;
        Roles r{};
        r.state = 0;
        if (les) {
            r.flags = K_WALL | K_EDDY;
            run(KN_WALL_EDDY, f, r);
            r.flags = K_FUSE_LES;
        }
        run(KN_VEL_PRECOMP, f, r);
        for (int k = 0; k < V; ++k) {
            FieldsT<T> g = f;
            for (int a = 0; a < 3; ++a) { g.vel[1][a] = tp.U[k][a]; g.vel[2][a] = tp.U[k + 1][a]; }
            Roles q{};
            q.state = 0; q.vin = 1; q.vout = 2;
            run(KN_VEL_SWEEP, g, q);
        }
        const V3p& ustar = tp.U[V];
        {
            FieldsT<T> g = f;
            for (int a = 0; a < 3; ++a) g.vel[1][a] = ustar[a];
            Roles q{};
            q.vout = 1;
            run(KN_PRES_PRECOMP, g, q);
        }
        record_pres(f, tp, 0);
        for (int a = 0; a < 3; ++a)
            if (int rc = copy(out_u[a], ustar[a])) return rc;
        if (int rc = copy(out_p, tp.P0)) return rc;
        {
            FieldsT<T> g = f;
            g.p = out_p;
            g.pc[0] = pc_final(tp, 0);
            for (int a = 0; a < 3; ++a) { g.vel[1][a] = out_u[a]; g.ucorr[a] = tp.UC[0][a]; }
            Roles q{};
            q.pin = 0; q.vout = 1;
            run(KN_APPLY_PRES, g, q);
            run(KN_VEL_CORRECT, g, q);
        }
        for (int c = 1; c <= C; ++c) {
            FieldsT<T> g = f;
            for (int a = 0; a < 3; ++a) { g.ucorr[a] = tp.UC[c - 1][a]; g.aodu[a] = tp.AO[c][a]; }
            g.invDpiso = tp.IDPI[c];
            g.bpiso = tp.BPI[c];
            Roles q{};
            q.state = 0;
            run(KN_PISO_AODU, g, q);
            run(KN_PISO_PRECOMP, g, q);
            record_pres(g, tp, c);
            g.p = out_p;
            g.pc[0] = pc_final(tp, c);
            for (int a = 0; a < 3; ++a) { g.vel[1][a] = out_u[a]; g.ucorr[a] = tp.UC[c][a]; }
            Roles w{};
            w.state = 0; w.pin = 0; w.vout = 1;
            run(KN_PISO_CORRECT, g, w);
        }
        return check_launch();
    }

    // ---- reverse of the one-shot kernels ----
    struct Seed {
        Dual<A>* arr;
        A* lam;
        int kind;   // StencilKind of the outputs one input cell reaches
        int face;   // -1: a cell field; else the boundary texture of this face
    };
    template <class Run>
    void vjp(Run&& go, const Pairs<A>& pr, const std::vector<Seed>& seeds, bool cs) {
        for (const Seed& sd : seeds) {
            const bool bnd = sd.face >= 0;
            const Colour col = bnd ? colour_bnd(sd.kind) : colour_of(sd.kind);
            const Stencil stn = make_stencil(sd.kind);
            const int W = bnd ? bw(sd.face) : X, H = bnd ? bh(sd.face) : Y, D = bnd ? 1 : Z;
            for (int k = 0; k < col.m; ++k) {
                dual_seed(sd.arr, W, H, D, col, k, st);
                go(pd);
                dual_contract(pr, tcon, N, st);
                if (bnd) adj_gather_bnd(tcon, X, Y, Z, sd.face, stn, col, k, sd.lam, st);
                else adj_gather(tcon, X, Y, Z, stn, col, k, sd.lam, st);
            }
            dual_seed(sd.arr, W, H, D, col, -1, st);
        }
        if (cs) {
            ParamsT<Dual<A>> q = pd;
            q.cs.d = (A)1;
            go(q);
            dual_contract(pr, tcon, N, st);
            adj_sum(tcon, N, lamCs, sum_part, st);
        }
    }
    void pack(const T* src, Dual<A>* dst) { dual_pack<T, A>(src, dst, N, st); }
    void pack3(const V3p& src, Dual<A>* const dst[3]) {
        for (int a = 0; a < 3; ++a) pack(src[a], dst[a]);
    }
    void dlaunch(Kernel k, const ParamsT<Dual<A>>& q, const Roles& r) { launch<Dual<A>>(k, q, fd, r, st); }

    // transposed pressure solve from the adjoint of its last iterate in
    // lamPC[0]: sweeps P-1 .. 0 in windows of ADJ_WIN; within a window each
    // sweep's input adjoint goes to a fresh slot, then one coefficient pass
    // covers the window; the last slot carries over to the next window
    void reverse_pres(const Tape& tp, int solve) {
        AdjPresSweep<T, A> a{};
        a.X = X; a.Y = Y; a.Z = Z;
        a.flag = s->f.flag;
        if (solve == 0) {
            a.cE = tp.cp[0]; a.invDp = tp.invDp; a.rhs = tp.bp;
            a.lam_cE = lamCP[0]; a.lam_invDp = lamInvDp; a.lam_rhs = lamBp;
        } else {
            a.cE = tp.cpe; a.invDp = tp.IDPI[solve]; a.rhs = tp.BPI[solve];
            a.lam_cE = lamCPE; a.lam_invDp = lamIDPI; a.lam_rhs = lamBPI;
        }
        a.cN = tp.cp[1]; a.cU = tp.cp[2];
        a.lam_cN = lamCP[1]; a.lam_cU = lamCP[2];
        const A* cur = lamPC[0];
        int ci = ADJ_WIN;  // slot of cur (lamPC[0] stands outside the slots)
        for (int k = P - 1; k >= 0;) {
            SweepWindow<T, A> w{};
            w.n = std::min(ADJ_WIN, k + 1);
            for (int j = 0; j < w.n; ++j, --k) {
                const T* xin = k > 0 ? tp.PS[solve][k - 1] : nullptr;
                w.lo[j][0] = cur;
                w.in[j][0] = xin;
                if (!xin) continue;  // the zero first iterate: its adjoint is not needed
                const int slot = (ci + 1 + j) % (ADJ_WIN + 1);
                a.lam_out = cur;
                a.lam_in = pwin[slot];
                adj_pres_lam(a, N, st);
                cur = pwin[slot];
            }
            adj_pres_coef(a, w, N, st);
            ci = (ci + w.n) % (ADJ_WIN + 1);
        }
    }

    // fixed-value faces are the only boundary textures the step reads
    std::vector<int> fixed_faces() const {
        std::vector<int> r;
        for (int f = 0; f < 6; ++f)
            if (s->p.velBnd[f] == UWS_VEL_FIXED_VALUE) r.push_back(f);
        return r;
    }

    Op1Args<A> op1_args() {
        Op1Args<A> a{};
        a.X = X; a.Y = Y; a.Z = Z; a.N = N;
        a.slots = s->cfg.wall_slots;
        a.les = les;
        int k = 0;
        for (int c = 0; c < 3; ++c) {
            a.lamS[c] = lamS[c]; a.lamG[c] = lamG[c]; a.lamD[c] = lamD[c];
            a.lamInvD[c] = lamInvD[c]; a.lamB[c] = lamB[c];
            a.lamU[c] = lamU0[c];
        }
        for (int d = 0; d < 6; ++d) a.lvis[d] = op1s[k++];
        for (int d = 0; d < 18; ++d) a.lfg[d] = op1s[k++];
        for (int d = 0; d < 6; ++d) a.ladv[d] = op1s[k++];
        for (int d = 0; d < 6; ++d) a.lpf[d] = op1s[k++];
        for (int c = 0; c < 3; ++c) a.lu[c] = op1s[k++];
        a.lp = op1s[k++];
        for (int c = 0; c < 9; ++c) a.lg[c] = op1s[k++];
        a.lnut = op1s[k++];
        a.lcs = op1s[k++];
        a.lamP = lamP;
        for (int f : fixed_faces())
            for (int c = 0; c < 3; ++c) a.lamBnd[f][c] = lamBnd[f][c];
        return a;
    }

    int reverse_step(const Tape& tp) {
        const uint8_t* flag = s->f.flag;
        zero_all({lamInvD[0], lamInvD[1], lamInvD[2], lamD[0], lamD[1], lamD[2], lamS[0], lamS[1], lamS[2],
              lamG[0], lamG[1], lamG[2], lamB[0], lamB[1], lamB[2], lamCP[0], lamCP[1], lamCP[2],
              lamU0[0], lamU0[1], lamU0[2], lamUc[0], lamUc[1], lamUc[2], lamCPE, lamInvDp, lamBp});
        V3a& lu = lamU[cur];
        // ---- PISO correctors, last first ----
        for (int c = C; c >= 1; --c) {
            // u += u'(p', 1/D, aodu), p += p' (fluid cells)
            for (int a = 0; a < 3; ++a) adj_add_fluid(flag, lamUc[a], lu[a], lamUT[a], N, st);
            adj_add_fluid<A>(flag, nullptr, lamP, lamPC[0], N, st);
            // this corrector's accumulators (lamUcNew is first written by the
            // aodu reverse below; its old contents are the previous lamUc)
            zero_all({lamAO[0], lamAO[1], lamAO[2], lamIDPI, lamBPI, lamUcNew[0], lamUcNew[1], lamUcNew[2]});
            if (hand) {
                AdjCorrect<T, A> a{};
                a.piso = true;
                a.flag = flag;
                a.pc = pc_final(tp, c);
                a.lam_pc = lamPC[0];
                for (int k = 0; k < 3; ++k) {
                    a.coef[k] = tp.invD[k]; a.ao[k] = tp.AO[c][k]; a.lam_uc[k] = lamUT[k];
                    a.lam_coef[k] = lamInvD[k]; a.lam_ao[k] = lamAO[k];
                }
                adj_correct(pd, a, st);
            } else {
                pack(pc_final(tp, c), fd.pc[0]);
                pack3(tp.invD, fd.invD);
                pack3(tp.AO[c], fd.aodu);
                Pairs<A> pr{};
                for (int a = 0; a < 3; ++a) { pr.lam[pr.n] = lamUT[a]; pr.d[pr.n++] = fd.ucorr[a]; }
                std::vector<Seed> sd{{fd.pc[0], lamPC[0], ST_STAR7, -1}};
                for (int a = 0; a < 3; ++a) sd.push_back({fd.invD[a], lamInvD[a], ST_OWN, -1});
                for (int a = 0; a < 3; ++a) sd.push_back({fd.aodu[a], lamAO[a], ST_OWN, -1});
                Roles r{};
                r.state = 0; r.pin = 0; r.vout = 1; r.flags = K_NO_INPLACE;
                vjp([&](const ParamsT<Dual<A>>& q) { dlaunch(KN_PISO_CORRECT, q, r); }, pr, sd, false);
            }
            reverse_pres(tp, c);
            // PISO coefficients (1/diag, rhs) of this corrector
            if (hand) {
                AdjPisoCoef<T, A> a{};
                a.flag = flag;
                a.cpe = tp.cpe;
                a.lam_idp = lamIDPI; a.lam_bp = lamBPI;
                a.lcf = op1s[0];
                a.lam_cpe = lamCPE;
                for (int k = 0; k < 3; ++k) {
                    a.invD[k] = tp.invD[k]; a.ao[k] = tp.AO[c][k]; a.cp[k] = tp.cp[k];
                    a.lam_invD[k] = lamInvD[k]; a.lam_ao[k] = lamAO[k]; a.lam_cp[k] = lamCP[k];
                }
                adj_piso_coef(pd, a, st);
            } else {
                pack(tp.cpe, fd.cpisoE);
                pack3(tp.cp, fd.cp);
                Pairs<A> pr{};
                pr.lam[pr.n] = lamIDPI; pr.d[pr.n++] = fd.invDpiso;
                pr.lam[pr.n] = lamBPI; pr.d[pr.n++] = fd.bpiso;
                std::vector<Seed> sd;
                for (int a = 0; a < 3; ++a) sd.push_back({fd.invD[a], lamInvD[a], ST_STAR7, -1});
                for (int a = 0; a < 3; ++a) sd.push_back({fd.aodu[a], lamAO[a], ST_STAR7, -1});
                sd.push_back({fd.cpisoE, lamCPE, ST_STAR7, -1});
                sd.push_back({fd.cp[1], lamCP[1], ST_STAR7, -1});
                sd.push_back({fd.cp[2], lamCP[2], ST_STAR7, -1});
                Roles r{};
                vjp([&](const ParamsT<Dual<A>>& q) { dlaunch(KN_PISO_PRECOMP, q, r); }, pr, sd, false);
            }
            // aodu(u^n, u' of the previous correction)
            if (hand) {
                AdjAodu<T, A> a{};
                a.flag = flag;
                for (int k = 0; k < 3; ++k) {
                    a.u[k] = tp.U[0][k]; a.uc[k] = tp.UC[c - 1][k]; a.lam_ao[k] = lamAO[k];
                    a.lam_u[k] = lamU0[k]; a.lam_uc[k] = lamUcNew[k];
                }
                adj_aodu(pd, a, st);
            } else {
                pack3(tp.U[0], fd.vel[0]);
                pack3(tp.UC[c - 1], fd.ucorr);
                Pairs<A> pr{};
                for (int a = 0; a < 3; ++a) { pr.lam[pr.n] = lamAO[a]; pr.d[pr.n++] = fd.aodu[a]; }
                std::vector<Seed> sd;
                for (int a = 0; a < 3; ++a) sd.push_back({fd.vel[0][a], lamU0[a], ST_STAR7, -1});
                for (int a = 0; a < 3; ++a) sd.push_back({fd.ucorr[a], lamUcNew[a], ST_STAR7, -1});
                Roles r{};
                r.state = 0;
                vjp([&](const ParamsT<Dual<A>>& q) { dlaunch(KN_PISO_AODU, q, r); }, pr, sd, false);
            }
            std::swap(lamUc, lamUcNew);
        }
        // ---- SIMPLE correction: u = u* + u'(p', D), p = p^n + p' ----
        for (int a = 0; a < 3; ++a) adj_add_fluid(flag, lamUc[a], lu[a], lamUT[a], N, st);
        // p = p^n + p' (fluid cells): lamPC[0] starts as lamP there, 0 elsewhere,
        // and the correction's reverse adds to it
        adj_add_fluid<A>(flag, nullptr, lamP, lamPC[0], N, st);
        if (hand) {
            AdjCorrect<T, A> a{};
            a.piso = false;
            a.flag = flag;
            a.pc = pc_final(tp, 0);
            a.lam_pc = lamPC[0];
            for (int k = 0; k < 3; ++k) {
                a.coef[k] = tp.D[k]; a.lam_uc[k] = lamUT[k]; a.lam_coef[k] = lamD[k];
            }
            adj_correct(pd, a, st);
        } else {
            pack(pc_final(tp, 0), fd.pc[0]);
            pack3(tp.D, fd.D);
            Pairs<A> pr{};
            for (int a = 0; a < 3; ++a) { pr.lam[pr.n] = lamUT[a]; pr.d[pr.n++] = fd.ucorr[a]; }
            std::vector<Seed> sd{{fd.pc[0], lamPC[0], ST_STAR7, -1}};
            for (int a = 0; a < 3; ++a) sd.push_back({fd.D[a], lamD[a], ST_OWN, -1});
            Roles r{};
            r.pin = 0; r.vout = 1; r.flags = K_NO_INPLACE;
            vjp([&](const ParamsT<Dual<A>>& q) { dlaunch(KN_VEL_CORRECT, q, r); }, pr, sd, false);
        }
        reverse_pres(tp, 0);
        // ---- Rhie-Chow coefficients ----
        if (hand) {
            AdjRhieChow<T, A> a{};
            a.flag = flag;
            a.p = tp.P0;
            a.lam_cpe = lamCPE; a.lam_idp = lamInvDp; a.lam_bp = lamBp;
            a.lam_p = lamP;
            for (int k = 0; k < 3; ++k) {
                a.invD[k] = tp.invD[k]; a.D[k] = tp.D[k]; a.lam_cp[k] = lamCP[k];
                a.lam_u[k] = lu[k]; a.lam_invD[k] = lamInvD[k]; a.lam_D[k] = lamD[k];
            }
            for (int d = 0; d < 6; ++d) { a.w[d] = op1s[d]; a.ldi[d] = op1s[6 + d]; }
            for (int f : fixed_faces())
                for (int k = 0; k < 3; ++k) a.lam_bnd[f][k] = lamBnd[f][k];
            adj_rhie_chow(pd, a, st);
        } else {
            pack3(tp.U[V], fd.vel[1]);
            pack(tp.P0, fd.p);
            pack3(tp.invD, fd.invD);
            Pairs<A> pr{};
            for (int a = 0; a < 3; ++a) { pr.lam[pr.n] = lamCP[a]; pr.d[pr.n++] = fd.cp[a]; }
            pr.lam[pr.n] = lamCPE; pr.d[pr.n++] = fd.cpisoE;
            pr.lam[pr.n] = lamInvDp; pr.d[pr.n++] = fd.invDp;
            pr.lam[pr.n] = lamBp; pr.d[pr.n++] = fd.bp;
            std::vector<Seed> sd;
            for (int a = 0; a < 3; ++a) sd.push_back({fd.vel[1][a], lu[a], ST_STAR7, -1});
            sd.push_back({fd.p, lamP, ST_AXIS13, -1});
            for (int a = 0; a < 3; ++a) sd.push_back({fd.invD[a], lamInvD[a], ST_STAR7, -1});
            for (int a = 0; a < 3; ++a) sd.push_back({fd.D[a], lamD[a], ST_OWN, -1});
            for (int f : fixed_faces())
                for (int a = 0; a < 3; ++a) sd.push_back({fd.bnd[f][a], lamBnd[f][a], ST_OWN, f});
            Roles r{};
            r.vout = 1;
            vjp([&](const ParamsT<Dual<A>>& q) { dlaunch(KN_PRES_PRECOMP, q, r); }, pr, sd, false);
        }
        // ---- momentum sweeps, transposed ----
        {
            AdjVelSweep<T, A> a{};
            a.X = X; a.Y = Y; a.Z = Z;
            a.q = (A)(s->p.ds * (T)0.25f);  // the forward's own coefficient
            a.flag = flag;
            for (int c = 0; c < 3; ++c) {
                a.S[c] = tp.S[c]; a.G[c] = tp.G[c]; a.invD[c] = tp.invD[c]; a.b[c] = tp.b[c];
                a.lam_S[c] = lamS[c]; a.lam_G[c] = lamG[c]; a.lam_invD[c] = lamInvD[c]; a.lam_b[c] = lamB[c];
            }
            // sweeps V-1 .. 0 in windows, as for the pressure; the adjoint of
            // u^0 (the last sweep's input) goes to lamU[1 - cur]
            const V3a* lo = &lamU[cur];
            int ci = ADJ_WIN;
            for (int k = V - 1; k >= 0;) {
                SweepWindow<T, A> w{};
                w.n = std::min(ADJ_WIN, k + 1);
                for (int j = 0; j < w.n; ++j, --k) {
                    const V3a* li = k > 0 ? &vwin[(ci + 1 + j) % (ADJ_WIN + 1)] : &lamU[1 - cur];
                    for (int c = 0; c < 3; ++c) {
                        w.lo[j][c] = (*lo)[c];
                        w.in[j][c] = tp.U[k][c];
                        a.lam_out[c] = (*lo)[c];
                        a.lam_in[c] = (*li)[c];
                    }
                    adj_vel_lam(a, N, st);
                    lo = li;
                }
                adj_vel_coef(a, w, N, st);
                ci = (ci + w.n) % (ADJ_WIN + 1);
            }
            if (V > 0) cur = 1 - cur;
        }
        // lamU0 += adjoint of u^0 from the sweeps (every cell)
        for (int a = 0; a < 3; ++a) adj_add_fluid<A>(nullptr, lamU0[a], lamU[cur][a], lamU0[a], N, st);
        // ---- Op1: wall function, eddy viscosity, LES term, momentum coefficients ----
        pack3(tp.U[0], fd.vel[0]);
        pack(tp.P0, fd.p);
        if (hand) {
            // the primal values in the adjoint arithmetic: the forward bodies
            // run once on the Dual mirror, every tangent zero
            Roles r{};
            r.state = 0;
            if (les) {
                r.flags = K_WALL | K_EDDY;
                dlaunch(KN_WALL_EDDY, pd, r);
                r.flags = K_FUSE_LES;
            }
            dlaunch(KN_VEL_PRECOMP, pd, r);
            const Op1Args<A> oa = op1_args();
            adj_op1<A>(pd, fd, oa, st);
            if (les) adj_sum(oa.lcs, N, lamCs, sum_part, st);
        } else {
            Pairs<A> pr{};
            for (int a = 0; a < 3; ++a) {
                pr.lam[pr.n] = lamS[a]; pr.d[pr.n++] = fd.S[a];
                pr.lam[pr.n] = lamG[a]; pr.d[pr.n++] = fd.G[a];
                pr.lam[pr.n] = lamD[a]; pr.d[pr.n++] = fd.D[a];
                pr.lam[pr.n] = lamInvD[a]; pr.d[pr.n++] = fd.invD[a];
                pr.lam[pr.n] = lamB[a]; pr.d[pr.n++] = fd.b[a];
            }
            std::vector<Seed> sd;
            for (int a = 0; a < 3; ++a) sd.push_back({fd.vel[0][a], lamU0[a], les ? ST_DIAMOND25 : ST_STAR7, -1});
            sd.push_back({fd.p, lamP, ST_STAR7, -1});
            for (int f : fixed_faces())
                for (int a = 0; a < 3; ++a) sd.push_back({fd.bnd[f][a], lamBnd[f][a], les ? ST_STAR7 : ST_OWN, f});
            const bool use_les = les;
            vjp(
                [&](const ParamsT<Dual<A>>& q) {
                    Roles r{};
                    r.state = 0;
                    if (use_les) {
                        r.flags = K_WALL | K_EDDY;
                        dlaunch(KN_WALL_EDDY, q, r);
                        r.flags = K_FUSE_LES;
                    }
                    dlaunch(KN_VEL_PRECOMP, q, r);
                },
                pr, sd, les);
        }
        // adjoint of u^n
        for (int a = 0; a < 3; ++a)
            if (int rc = copy(lamU[cur][a], lamU0[a])) return rc;
        return check_launch();
    }

    int backward(const double* lam_u, int64_t n_u, const double* lam_p, int64_t n_p) override {
        if (!fwd_done || !lam_u || n_u != 3LL * N || (lam_p && n_p != N)) return UWS_ERR_INVALID_ARG;
        for (int a = 0; a < 3; ++a) {
            std::vector<A> plane((size_t)N);
            for (int i = 0; i < N; ++i) plane[i] = lam_u[(size_t)i * 3 + a];
            if (hipMemcpyAsync(lamU[0][a], plane.data(), (size_t)N * sizeof(A), hipMemcpyHostToDevice, st) !=
                hipSuccess)
                return UWS_ERR_DEVICE;
            if (int rc = s->sync()) return rc;
        }
        if (lam_p) {
            if (hipMemcpyAsync(lamP, lam_p, (size_t)N * sizeof(A), hipMemcpyHostToDevice, st) != hipSuccess)
                return UWS_ERR_DEVICE;
            if (int rc = s->sync()) return rc;
        } else if (int rc = clear(lamP)) {
            return rc;
        }
        set_dual_params();
        cur = 0;
        for (int f = 0; f < 6; ++f)
            for (int a = 0; a < 3; ++a)
                if (int rc = clear(lamBnd[f][a], (size_t)bw(f) * bh(f))) return rc;
        if (hipMemsetAsync(lamCs, 0, sizeof(double), st) != hipSuccess) return UWS_ERR_DEVICE;
        for (int f = 0; f < 6; ++f)
            for (int a = 0; a < 3; ++a) dual_pack<T, A>(s->f.bnd[f][a], fd.bnd[f][a], bw(f) * bh(f), st);
        for (int seg = nseg - 1; seg >= 0; --seg) {
            const int n0 = seg * K, ns = std::min(nsteps, n0 + K) - n0;
            Tape& t0 = tapes[0];
            for (int a = 0; a < 3; ++a)
                if (int rc = copy(t0.U[0][a], ckpt[seg][a])) return rc;
            if (int rc = copy(t0.P0, ckpt[seg][3])) return rc;
            for (int j = 0; j < ns; ++j) {
                const bool last = j + 1 == ns;
                if (int rc = record_step(tapes[j], last ? seg_u : tapes[j + 1].U[0], last ? seg_p : tapes[j + 1].P0))
                    return rc;
            }
            for (int j = ns - 1; j >= 0; --j)
                if (int rc = reverse_step(tapes[j])) return rc;
            if (int rc = s->sync()) return rc;
        }
        have_grad = true;
        return UWS_OK;
    }

    int get_grad(int what, int face, double* dst, int64_t n) override {
        if (!have_grad || !dst) return UWS_ERR_INVALID_ARG;
        auto down = [&](const A* src, size_t np, int nc, int c) -> int {
            std::vector<A> tmp(np);
            if (hipMemcpyAsync(tmp.data(), src, np * sizeof(A), hipMemcpyDeviceToHost, st) != hipSuccess)
                return UWS_ERR_DEVICE;
            if (int rc = s->sync()) return rc;
            for (size_t i = 0; i < np; ++i) dst[i * nc + c] = (double)tmp[i];
            return UWS_OK;
        };
        switch (what) {
        case UWS_GRAD_VEL0:
            if (n != 3LL * N) return UWS_ERR_INVALID_ARG;
            for (int a = 0; a < 3; ++a)
                if (int rc = down(lamU[cur][a], N, 3, a)) return rc;
            return UWS_OK;
        case UWS_GRAD_PRES0:
            if (n != N) return UWS_ERR_INVALID_ARG;
            return down(lamP, N, 1, 0);
        case UWS_GRAD_BND: {
            if (face < 0 || face > 5) return UWS_ERR_INVALID_ARG;
            const size_t np = (size_t)bw(face) * bh(face);
            if (n != (int64_t)(3 * np)) return UWS_ERR_INVALID_ARG;
            for (int a = 0; a < 3; ++a)
                if (int rc = down(lamBnd[face][a], np, 3, a)) return rc;
            return UWS_OK;
        }
        case UWS_GRAD_CS:
            if (n != 1) return UWS_ERR_INVALID_ARG;
            if (hipMemcpyAsync(dst, lamCs, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
                return UWS_ERR_DEVICE;
            return s->sync();
        default: return UWS_ERR_INVALID_ARG;
        }
    }
};

}  // namespace uws

struct uws_adjoint {
    std::unique_ptr<uws::AdjointBase> impl;
};

using namespace uws;

namespace {
template <class F>
int with_adjoint(uws_adjoint_handle a, F&& fn) {
    if (!a || !a->impl) return UWS_ERR_INVALID_ARG;
    if (hipSetDevice(a->impl->device) != hipSuccess) return UWS_ERR_DEVICE;
    return fn(*a->impl);
}
template <class T>
int make_adjoint(SolverBase* base, int nsteps, int K, std::unique_ptr<AdjointBase>* out) {
    std::unique_ptr<Adjoint<T>> a(new (std::nothrow) Adjoint<T>());
    if (!a) return UWS_ERR_ALLOC;
    if (int rc = a->init(static_cast<Solver<T>*>(base), nsteps, K)) return rc;
    *out = std::move(a);
    return UWS_OK;
}
}  // namespace

extern "C" {

int uws_adjoint_create(uws_handle h, int32_t nsteps, int32_t ckpt_every, uws_adjoint_handle* out) {
    if (!h || !h->impl || !out || nsteps < 1 || ckpt_every < 1) return UWS_ERR_INVALID_ARG;
    *out = nullptr;
    SolverBase* s = h->impl.get();
    if (hipSetDevice(s->device) != hipSuccess) return UWS_ERR_DEVICE;
    const int prec = s->cfg.precision;
    if (prec != UWS_PREC_FP32 && prec != UWS_PREC_FP64) return UWS_ERR_UNSUPPORTED;
    if (s->sc.cal_residual || s->sc.lowp || s->sc.normalise || s->nonuniform()) return UWS_ERR_UNSUPPORTED;
    std::unique_ptr<AdjointBase> impl;
    const int rc = prec == UWS_PREC_FP64 ? make_adjoint<double>(s, nsteps, ckpt_every, &impl)
                                         : make_adjoint<float>(s, nsteps, ckpt_every, &impl);
    if (rc) return rc;
    uws_adjoint* a = new (std::nothrow) uws_adjoint();
    if (!a) return UWS_ERR_ALLOC;
    a->impl = std::move(impl);
    *out = a;
    return UWS_OK;
}

int uws_adjoint_destroy(uws_adjoint_handle a) {
    if (!a) return UWS_ERR_INVALID_ARG;
    if (a->impl) (void)hipSetDevice(a->impl->device);
    delete a;
    return UWS_OK;
}

int uws_adjoint_forward(uws_adjoint_handle a) {
    return with_adjoint(a, [&](AdjointBase& b) { return b.forward(); });
}

int uws_adjoint_backward(uws_adjoint_handle a, const double* lam_u, int64_t n_u, const double* lam_p, int64_t n_p) {
    return with_adjoint(a, [&](AdjointBase& b) { return b.backward(lam_u, n_u, lam_p, n_p); });
}

int uws_adjoint_get_grad(uws_adjoint_handle a, int32_t what, int32_t face, double* dst, int64_t n) {
    return with_adjoint(a, [&](AdjointBase& b) { return b.get_grad(what, face, dst, n); });
}

int uws_adjoint_memory(uws_adjoint_handle a, int64_t* bytes) {
    if (!bytes) return UWS_ERR_INVALID_ARG;
    return with_adjoint(a, [&](AdjointBase& b) {
        *bytes = (int64_t)b.bytes;
        return (int)UWS_OK;
    });
}

}  // extern "C"
