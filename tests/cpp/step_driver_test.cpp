// The step driver (uws_host.h: enqueue_step, stage_vel_sweeps, pres_solve)
// run against a recording launcher on the host: every buffer carries the
// number of Jacobi sweeps its contents have seen, and each launch checks and
// updates those counts.  Checked for every combination of multi-sweep
// kernels (pressure pairs, triples and four-sweep launches, momentum pairs),
// sweep counts, norms, PISO correctors and low-precision rounds:
//  - no launch reads and writes the same buffer, and u^n is never written
//    by the momentum sweeps;
//  - the momentum result holds V sweeps of u^n, every pressure solve P
//    sweeps (R rounds of K in the low-precision mode) when a correction
//    reads it;
//  - the residual kernels see the last sweep's input and output (a single
//    sweep), as the FP64 norms require;
//  - without a residual-based exit the new state buffer is advance_roles',
//    whether or not the momentum sweeps run in pairs.
// Exit status 0 when every case passes; the first failure is printed.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "uws_host.h"

using namespace uws;

namespace {

int g_fail = 0;
std::string g_case;

#define CHECK(cond)                                                                             \
    do {                                                                                        \
        if (!(cond)) {                                                                          \
            if (!g_fail) std::printf("FAIL [%s] line %d: %s\n", g_case.c_str(), __LINE__, #cond); \
            g_fail = 1;                                                                         \
        }                                                                                       \
    } while (0)

constexpr int GARBAGE = -1000;

struct Rec {
    bool f2 = false, fv = false, f3 = false, f4 = false;  // pressure pairs, momentum pairs, triples, fours
    int V = 0, P = 0, K = 0;
    int state = 0;
    int vel[3] = {GARBAGE, GARBAGE, GARBAGE};  // sweeps seen by each velocity buffer
    int pc[2] = {GARBAGE, GARBAGE};           // sweeps seen by each p' buffer
    int e[2] = {GARBAGE, GARBAGE};            // sweeps seen by each 2-byte e buffer (this round)
    int vel_sweeps_launched = 0;
    int last_vel_kind = 0;                    // 1 single, 2 pair
    int pres_in_solve = 0;                    // sweeps so far in the current solve
    int last_pres_kind = 0;
    int rounds_done = 0;

    int pv(int i) const { return i < 0 ? 0 : pc[i]; }
    void operator()(Kernel k, const Roles& r, int = 0) {
        switch (k) {
        case KN_VEL_SWEEP:
        case KN_VEL_SWEEP2: {
            const int n = k == KN_VEL_SWEEP2 ? 2 : 1;
            if (n == 2) CHECK(fv);
            CHECK(r.vin != r.vout);
            CHECK(r.vout != state);  // u^n stays intact
            CHECK(vel[r.vin] >= 0);
            vel[r.vout] = vel[r.vin] + n;
            vel_sweeps_launched += n;
            last_vel_kind = n;
            break;
        }
        case KN_RES_VEL:
            // residual of the last sweep: its input and output buffers
            CHECK(last_vel_kind == 1);
            CHECK(vel[r.vout] == vel[r.vin] + 1);
            break;
        case KN_PRES_SWEEP:
        case KN_PRES_SWEEP2:
        case KN_PRES_SWEEP3:
        case KN_PRES_SWEEP4: {
            const int n = k == KN_PRES_SWEEP4 ? 4 : (k == KN_PRES_SWEEP3 ? 3 : (k == KN_PRES_SWEEP2 ? 2 : 1));
            if (n == 2) CHECK(f2);
            if (n == 3) CHECK(f3);
            if (n == 4) CHECK(f4);
            CHECK(r.pin != r.pout);
            if (r.pin < 0) CHECK(pres_in_solve == 0);
            else CHECK(pc[r.pin] == pres_in_solve);
            pc[r.pout] = pv(r.pin) + n;
            pres_in_solve += n;
            last_pres_kind = n;
            break;
        }
        case KN_RES_PRES:
            CHECK(last_pres_kind == 1);
            CHECK(pc[r.pout] == pv(r.pin) + 1);
            break;
        case KN_DEFECT_RESIDUAL:
            // a round starts from the last round's p' (or zero)
            CHECK(pv(r.pin) == rounds_done * K);
            e[0] = e[1] = GARBAGE;
            break;
        case KN_LOWP_SWEEP:
        case KN_LOWP_SWEEP3:
        case KN_LOWP_SWEEP4: {
            const int n = k == KN_LOWP_SWEEP4 ? 4 : (k == KN_LOWP_SWEEP3 ? 3 : 1);
            if (n == 3) CHECK(f3);
            if (n == 4) CHECK(f4);
            const bool first = (r.flags & K_LOWP_FIRST) != 0, last = (r.flags & K_LOWP_LAST) != 0;
            int have = 0;
            if (!first) {
                CHECK(r.ein != r.eout);
                CHECK(e[r.ein] >= 0);
                have = e[r.ein];
            } else {
                CHECK(have == 0);
            }
            if (last) {
                CHECK(have + n == K);
                pc[r.pout] = pv(r.pin) + K;  // p' + e / s: one more round of K sweeps
                rounds_done += 1;
            } else {
                e[r.eout] = have + n;
            }
            break;
        }
        case KN_APPLY_PRES:
        case KN_VEL_CORRECT:
            // the SIMPLE correction reads a whole solve (R K sweeps in the low-precision mode)
            CHECK(pv(r.pin) == P);
            break;
        case KN_PISO_CORRECT:
            CHECK(pv(r.pin) == P);
            break;
        case KN_FILL_PC:
            pc[r.pout] = 0;
            break;
        default: break;
        }
        // a new solve starts after its coefficients are computed
        if (k == KN_PRES_PRECOMP || k == KN_PISO_PRECOMP) {
            pres_in_solve = 0;
            rounds_done = 0;
        }
    }
    int copy_vel(int dst, int src) {
        vel[dst] = vel[src];
        return UWS_OK;
    }
    int read_norm(int, double* out) {
        *out = 1.0;  // never below the tolerance: no early exit
        return UWS_OK;
    }
    double ncells() const { return 1000.0; }
    bool fused() const { return f2; }
    bool fused_vel() const { return fv; }
    bool fused3() const { return f3; }
    bool fused4() const { return f4; }
};

void run_case(int V, int P, int C, bool norms, int fl, bool lowp, int R, int K, bool residual, int state0) {
    // fl: bit 0 pressure pairs, 1 momentum pairs, 2 triples, 3 four-sweep launches
    const bool f2 = fl & 1, fv = (fl >> 1) & 1, f3 = (fl >> 2) & 1, f4 = (fl >> 3) & 1;
    char name[220];
    std::snprintf(name, sizeof name, "V=%d P=%d C=%d norms=%d pairs=%d vel_pairs=%d triples=%d fours=%d lowp=%d R=%d "
                                     "K=%d residual=%d state=%d",
                  V, P, C, (int)norms, (int)f2, (int)fv, (int)f3, (int)f4, (int)lowp, R, K, (int)residual, state0);
    g_case = name;
    StepCfg c;
    c.V = V; c.P = lowp ? R * K : P; c.C = C; c.piso = C > 0; c.norms = norms; c.lowp = lowp; c.R = R; c.K = K;
    c.cal_residual = residual; c.vci = 1; c.pci = 1; c.vtol = 1e-30f; c.ptol = 1e-30f;
    Rec l;
    l.f2 = f2; l.fv = fv; l.f3 = f3; l.f4 = f4; l.V = V; l.P = c.P; l.K = K;
    l.state = state0;
    l.vel[state0] = 0;
    RoleState rs;
    rs.state = state0; rs.cur = -1; rs.pc_latest = -1; rs.piso_done = 0;
    const int rc = enqueue_step(l, c, rs);
    CHECK(rc == UWS_OK);
    CHECK(l.vel_sweeps_launched == V);
    // the new state holds u* after V sweeps (the corrections update it in place)
    CHECK(rs.state >= 0 && rs.state < 3);
    CHECK(l.vel[rs.state] == V);
    if (V > 0) CHECK(rs.state != state0);
    // without a residual-based exit the step's role change is advance_roles',
    // with and without momentum pairs (launch groups sharing roles may differ
    // in them)
    if (!residual) {
        RoleState a;
        a.state = state0; a.cur = -1; a.pc_latest = -1; a.piso_done = 0;
        advance_roles(c, a);
        CHECK(a.state == rs.state);
    }
}

}  // namespace

int main() {
    int ncases = 0;
    for (int state0 = 0; state0 < 3; ++state0)
        for (int V : {0, 1, 2, 3, 5, 6, 50})
            for (int P : {0, 1, 2, 3, 4, 5, 7, 50})
                for (int C : {0, 2})
                    for (int norms = 0; norms < 2; ++norms)
                        for (int fl = 0; fl < 16; ++fl) {
                            run_case(V, P, C, norms, fl, false, 0, 0, false, state0);
                            ++ncases;
                        }
    // low-precision rounds (FP32 momentum, 2-byte corrections)
    for (int R : {1, 3})
        for (int K : {1, 2, 3, 4, 5, 7, 8, 20})
            for (int fl = 0; fl < 16; ++fl) {
                run_case(6, 0, 2, true, fl, true, R, K, false, 0);
                ++ncases;
            }
    // residual-based exit: single sweeps only, whatever the launcher offers
    for (int fl = 0; fl < 16; ++fl) {
        run_case(7, 9, 0, true, fl, false, 0, 0, true, 1);
        ++ncases;
    }
    if (g_fail) return 1;
    std::printf("step driver: %d cases passed\n", ncases);
    return 0;
}
