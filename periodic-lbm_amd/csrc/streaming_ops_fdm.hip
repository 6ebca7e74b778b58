// streaming_ops_fdm.hip — the stencils of stream_fdm_bardow that the reference selects at
// compile time (src/fvm_bardow.F90:591-669: -DFDM_WLS, -DFDM_WLS_GAUSS_V1,
// -DFDM_WLS_GAUSS_V2, -DFDM_ISO), as streaming operators PLBM_STREAM_FDM_BARDOW_WLS ..
// PLBM_STREAM_FDM_BARDOW_ISO of one library: the same kernels as the default operators
// (fd_ops.hpp), instantiated here so that the two translation units compile in parallel.
#include "fd_ops.hpp"

namespace plbm {

int launch_fd_variant(plbm_lattice *h, int op, int model, const void *src, void *dst, double dt) {
    switch (op) {
    case PLBM_STREAM_FDM_BARDOW_WLS:
        return launch_fd_op<PLBM_STREAM_FDM_BARDOW_WLS>(h, model, src, dst, dt);
    case PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1:
        return launch_fd_op<PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1>(h, model, src, dst, dt);
    case PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2:
        return launch_fd_op<PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2>(h, model, src, dst, dt);
    case PLBM_STREAM_FDM_BARDOW_ISO:
        return launch_fd_op<PLBM_STREAM_FDM_BARDOW_ISO>(h, model, src, dst, dt);
    }
    set_error("unknown streaming operator " + std::to_string(op));
    return PLBM_ERR_ARG;
}

int launch_fd_multi_variant(plbm_lattice *h, int op, int model, int ns, const void *src,
                            void *dst, double dt, void *lag) {
    switch (op) {
    case PLBM_STREAM_FDM_BARDOW_WLS:
        return launch_fd_multi_op<PLBM_STREAM_FDM_BARDOW_WLS>(h, model, ns, src, dst, dt, lag);
    case PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1:
        return launch_fd_multi_op<PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V1>(h, model, ns, src, dst,
                                                                      dt, lag);
    case PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2:
        return launch_fd_multi_op<PLBM_STREAM_FDM_BARDOW_WLS_GAUSS_V2>(h, model, ns, src, dst,
                                                                      dt, lag);
    case PLBM_STREAM_FDM_BARDOW_ISO:
        return launch_fd_multi_op<PLBM_STREAM_FDM_BARDOW_ISO>(h, model, ns, src, dst, dt, lag);
    }
    set_error("unknown streaming operator " + std::to_string(op));
    return PLBM_ERR_ARG;
}

}  // namespace plbm
