// One translation unit per specialised axis length: compiled with
// -DQSM_PASS_N=<length> it defines launch_fixed_<length>(), with
// -DQSM_PASS_N=0 the runtime-plan fallback launch_generic().
#include "qsm_kernels.hpp"

#ifndef QSM_PASS_N
#error "compile with -DQSM_PASS_N=<axis length> (0 for the runtime-plan kernels)"
#endif
#if QSM_PASS_N == 0 && !defined(QSM_GENERIC_PART)
#error "the runtime-plan kernels compile with -DQSM_GENERIC_PART=0..3"
#endif

namespace qsm {

[[maybe_unused]] static bool is_row_kind(int kind) {
    return kind == P_R2C_TVDIV || kind == P_R2C_F64 || kind == P_R2C_U8 || kind == P_R2C_F32 ||
           kind == P_R2C_SPHERE || kind == P_C2R_TV || kind == P_C2R_VS_SELECT || kind == P_C2R_VS_OUT ||
           kind == P_R2C_TVDIV_REC || kind == P_R2C_TVADJ_FIRST || kind == P_R2C_TVADJ ||
           kind == P_R2C_ACC || kind == P_C2R_ADJ || kind == P_C2R_F64OUT || kind == P_C2R_F64_MASKED ||
           kind == P_C2R_X0 || kind == P_R2C_X0 || kind == P_R2C_WRAPLAP || kind == P_R2C_TVDIV_W ||
           kind == P_R2C_NLTV_BND;
}

template <int N, int EG, int RBS>
static void launch_rows_s(int kind, const LaunchTable& lt, const PassArgs& a) {
    const dim3 g(lt.start[lt.nvol]), b(kThreads);
    hipStream_t st = a.stream;
    switch (kind) {
        case P_R2C_TVDIV:
            k_r2c_x<N, EG, SRC_TVDIV, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.parity, a.status,
                                                       a.status_rw, a.mirror, 0);
            break;
        case P_R2C_TVDIV_REC:
            k_r2c_x<N, EG, SRC_TVDIV_REC, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.parity, a.status,
                                                           a.status_rw, a.mirror, a.ridx);
            break;
        case P_R2C_TVADJ_FIRST:
            k_r2c_x<N, EG, SRC_TVADJ_FIRST, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.parity, nullptr,
                                                             nullptr, nullptr, a.ridx);
            break;
        case P_R2C_TVADJ:
            k_r2c_x<N, EG, SRC_TVADJ, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.parity, nullptr, nullptr,
                                                       nullptr, a.ridx);
            break;
        case P_R2C_F64: k_r2c_x<N, EG, SRC_F64, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, a.status, nullptr, nullptr, 0); break;
        case P_R2C_U8: k_r2c_x<N, EG, SRC_U8, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, a.status, nullptr, nullptr, 0); break;
        case P_R2C_F32: k_r2c_x<N, EG, SRC_F32, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, a.status, nullptr, nullptr, 0); break;
        case P_R2C_ACC: k_r2c_x<N, EG, SRC_ACC, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, nullptr, nullptr, nullptr, 0); break;
        case P_R2C_SPHERE: k_r2c_x<N, EG, SRC_SPHERE, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.ridx, a.status, nullptr, nullptr, 0); break;
        case P_C2R_TV: k_c2r_tv<N, EG, C2R_TV, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.parity, a.status); break;
        case P_C2R_ADJ: k_c2r_tv<N, EG, C2R_ADJ, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, 0, nullptr); break;
        case P_C2R_F64OUT: k_c2r_tv<N, EG, C2R_F64, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, 0, nullptr); break;
        case P_C2R_F64_MASKED: k_c2r_tv<N, EG, C2R_F64_MASKED, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, 0, nullptr); break;
        case P_C2R_X0: k_c2r_tv<N, EG, C2R_X0, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, 0, nullptr); break;
        case P_R2C_X0: k_r2c_x<N, EG, SRC_X0, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, nullptr, nullptr, nullptr, 0); break;
        case P_R2C_TVDIV_W:
            k_r2c_x<N, EG, SRC_TVDIV_W, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.parity, a.status,
                                                         nullptr, nullptr, a.ridx);
            break;
        case P_R2C_NLTV_BND:
            k_r2c_x<N, EG, SRC_NLTV_BND, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.parity, nullptr,
                                                          a.status_rw, nullptr, a.ridx);
            break;
        case P_R2C_WRAPLAP: k_r2c_x<N, EG, SRC_WRAPLAP, RBS><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, nullptr, nullptr, nullptr, 0); break;
        case P_C2R_VS_SELECT: k_c2r_vs_select<N, EG><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.ridx, a.delta); break;
        case P_C2R_VS_OUT: k_c2r_vs_out<N, EG><<<g, b, 0, st>>>(a.vols, lt, a.tw); break;
        default: break;
    }
}

// xshift 2: the R2C / TV C2R kernels with a quarter of the rows per block
// (compiled-length kernels only; the runtime-plan ones read rows_x).
template <int N, int EG>
static void launch_rows(int kind, const LaunchTable& lt, const PassArgs& a) {
    if constexpr (N > 0) {
        if (a.xshift == 2 && (qsmfft::fixed_rows(N) >> 2) >= 4) {
            launch_rows_s<N, EG, 2>(kind, lt, a);
            return;
        }
    }
    launch_rows_s<N, EG, 0>(kind, lt, a);
}

template <int N, int EG, int TW = kTile>
static void launch_cols(int kind, const LaunchTable& lt, const PassArgs& a) {
    // lengths without the two-stage transform compile only the staged one
    constexpr int S0 = qsmfft::col2_ok(N) ? 0 : 1;
    const dim3 g(lt.start[lt.nvol]), b(kThreads);
    hipStream_t st = a.stream;
    switch (kind) {
        case P_Y_FWD:
            if (a.sched) k_y_pass<N, EG, false, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.sel2, a.status);
            else k_y_pass<N, EG, false, S0, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.sel2, a.status);
            break;
        case P_Y_INV:
            if (a.sched) k_y_pass<N, EG, true, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.sel2, a.status);
            else k_y_pass<N, EG, true, S0, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, a.sel2, a.status);
            break;
        case P_Z_TV_SETUP: k_z_tv<N, EG, ZOP_TV_SETUP, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status); break;
        case P_Z_TV_SOLVE:
            if (a.sched) k_z_tv<N, EG, ZOP_TV_SOLVE, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status);
            else k_z_tv<N, EG, ZOP_TV_SOLVE, S0, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status);
            break;
        case P_Z_TV_ADJ:
            if (a.sched) k_z_tv<N, EG, ZOP_TV_ADJ, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, nullptr);
            else k_z_tv<N, EG, ZOP_TV_ADJ, S0, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, nullptr);
            break;
        case P_Z_TV_APPLYA: k_z_tv<N, EG, ZOP_TV_APPLYA, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, nullptr); break;
        case P_Z_DIPOLE: k_z_tv<N, EG, ZOP_DIPOLE, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, nullptr); break;
        case P_Z_FILTER: k_z_tv<N, EG, ZOP_FILTER, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, nullptr); break;
        case P_Z_RTS_SETUP: k_z_tv<N, EG, ZOP_RTS_SETUP, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, nullptr); break;
        case P_Z_TV_SOLVE0:
            if (a.sched) k_z_tv<N, EG, ZOP_TV_SOLVE0, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status);
            else k_z_tv<N, EG, ZOP_TV_SOLVE0, S0, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status);
            break;
        case P_Z_RTS_SOLVE:
            if (a.sched) k_z_tv<N, EG, ZOP_RTS_SOLVE, 1, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status);
            else k_z_tv<N, EG, ZOP_RTS_SOLVE, S0, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.status);
            break;
        case P_Z_VS_FWD: k_z_vs<N, EG, ZOP_VS_FWD, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, a.sel, 0, 0); break;
        case P_Z_VS_RADIUS: k_z_vs<N, EG, ZOP_VS_RADIUS, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, 0, a.ridx, a.first); break;
        case P_Z_VS_DECONV: k_z_vs<N, EG, ZOP_VS_DECONV, TW><<<g, b, 0, st>>>(a.vols, lt, a.tw, 0, 0, a.divide); break;
        default: break;
    }
}

#if QSM_PASS_N > 0
#define QSM_CAT2(a, b) a##b
#define QSM_CAT(a, b) QSM_CAT2(a, b)
void QSM_CAT(launch_fixed_, QSM_PASS_N)(int kind, const LaunchTable& lt, const PassArgs& a) {
    if (is_row_kind(kind)) launch_rows<QSM_PASS_N, 0>(kind, lt, a);
    else launch_cols<QSM_PASS_N, 0>(kind, lt, a);
}
#elif QSM_GENERIC_PART == 0
// The runtime-plan kernels are compiled in four parts (this one with the row
// kernels and the dispatcher) so that they build in parallel.
void launch_generic_cols(int kind, const LaunchTable& lt, const PassArgs& a);
void launch_generic_cols_long16(int kind, const LaunchTable& lt, const PassArgs& a);
void launch_generic_cols_long32(int kind, const LaunchTable& lt, const PassArgs& a);
void launch_generic(int kind, int emax, const LaunchTable& lt, const PassArgs& a) {
    if (is_row_kind(kind)) launch_rows<0, 16>(kind, lt, a);
    else if (emax == kColLongTile + 16) launch_generic_cols_long16(kind, lt, a);  // work length <= 512
    else if (emax > kColLongTile) launch_generic_cols_long32(kind, lt, a);        // <= 1024
    else launch_generic_cols(kind, lt, a);                                        // <= 256
}
#elif QSM_GENERIC_PART == 1
void launch_generic_cols(int kind, const LaunchTable& lt, const PassArgs& a) { launch_cols<0, 16>(kind, lt, a); }
#elif QSM_GENERIC_PART == 2
void launch_generic_cols_long16(int kind, const LaunchTable& lt, const PassArgs& a) {
    launch_cols<0, 16, kTileLong>(kind, lt, a);
}
#else
void launch_generic_cols_long32(int kind, const LaunchTable& lt, const PassArgs& a) {
    launch_cols<0, 32, kTileLong>(kind, lt, a);
}
#endif

}  // namespace qsm
