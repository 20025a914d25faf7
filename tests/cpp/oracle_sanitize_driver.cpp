// The CPU oracle (oracle/uws_oracle.cpp, compiled in whole) built with
// AddressSanitizer and UBSan, replaying the cases a test writes to a file:
// the reference-style initialisation, every stage of a step once in order,
// two whole steps, then every field, boundary texture and (stretched grid)
// face flux read back.  Record layout, native endianness:
//   int32 precision (0 FP32, 1 FP64) | uws_config | int32 inflow type |
//   float wind speed | int32 box solid | float velZn[3] | float ref height |
//   float alpha | int32 flags[X*Y*Z] | int32 has_faces |
//   float faces[(X+1)+(Y+1)+(Z+1)] when has_faces
// Exit status 0 when every record ran; the sanitizers abort on the first
// error they find.
#include "../../oracle/uws_oracle.cpp"

#include <cstdio>
#include <vector>

namespace {

template <class T>
bool get(FILE* f, T* v, size_t n = 1) {
    return std::fread(v, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::printf("usage: %s cases.bin\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) {
        std::printf("cannot open %s\n", argv[1]);
        return 2;
    }
    int ran = 0;
    for (;;) {
        int32_t prec = 0;
        if (!get(f, &prec)) break;
        uws_config c;
        int32_t inflow = 0, box = 0, has_faces = 0;
        float wind = 0, velZn[3] = {0, 0, 0}, refh = 0, alpha = 0;
        if (!get(f, &c) || !get(f, &inflow) || !get(f, &wind) || !get(f, &box) || !get(f, velZn, 3) ||
            !get(f, &refh) || !get(f, &alpha)) {
            std::printf("truncated record %d\n", ran);
            return 2;
        }
        const int X = c.grid_res[0], Y = c.grid_res[1], Z = c.grid_res[2];
        const int64_t N = (int64_t)X * Y * Z;
        std::vector<int32_t> flags((size_t)N);
        if (!get(f, flags.data(), flags.size()) || !get(f, &has_faces)) {
            std::printf("truncated record %d\n", ran);
            return 2;
        }
        std::vector<float> faces;
        if (has_faces) {
            faces.resize((size_t)(X + Y + Z + 3));
            if (!get(f, faces.data(), faces.size())) {
                std::printf("truncated record %d\n", ran);
                return 2;
            }
        }
        void* h = uwso_create(&c, prec);
        if (has_faces && uwso_set_face_positions(h, faces.data(), faces.data() + X + 1, faces.data() + X + Y + 2) != 0) {
            std::printf("record %d: face positions refused\n", ran);
            return 1;
        }
        // inflow type 2 initialises the fields as type 0 at rest (the tests' oracle wrapper does the same)
        uwso_init_vel_pres_fields(h, inflow == 2 ? 0 : inflow, inflow == 2 ? 0.f : wind);
        if (box) uwso_init_box(h);
        uwso_set_fixed_value_vel_bnd_cond(h, inflow, wind, velZn, refh, alpha);
        uwso_init_flags(h, flags.data());
        const bool les = c.turbulence_model == 1, piso = c.fvm_solver_type == 1;
        for (int s = 0; s < UWS_STAGE_COUNT; ++s) {
            if (!les && (s == UWS_STAGE_WALL_FUNC || s == UWS_STAGE_EDDY_VIS || s == UWS_STAGE_LES_TERM)) continue;
            if (!piso && s == UWS_STAGE_PISO) continue;
            uwso_run_stage(h, s);
        }
        uwso_step(h, 2);
        std::vector<double> buf((size_t)(3 * (int64_t)(X + 1) * (Y + 1) * (Z + 1) + 9 * N));
        int fields = 0;
        for (int fid = 0; fid < 64; ++fid) fields += uwso_get_field(h, fid, buf.data()) == 0;
        for (int face = 0; face < 6; ++face) uwso_get_bnd_field(h, face, buf.data());
        if (has_faces) uwso_get_face_flux(h, buf.data());
        double res[2];
        int sweeps[2];
        uwso_residuals(h, res);
        uwso_sweep_counts(h, sweeps);
        uwso_destroy(h);
        std::printf("record %d: %dx%dx%d precision %d, %d fields read\n", ran, X, Y, Z, prec, fields);
        ++ran;
    }
    std::fclose(f);
    std::printf("%d cases ran\n", ran);
    return ran > 0 ? 0 : 1;
}
