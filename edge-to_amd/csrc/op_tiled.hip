// Launchers of the LDS-tiled fine-grid operator (see op_tiled.hpp).
#include "op_tiled.hpp"
#include "stiffness.hpp"

namespace et {

static int g_tiled_jc = 8;  // y planes per workgroup (tunable)

void set_tiled_jc(int jc) { g_tiled_jc = jc > 0 ? jc : 8; }
int get_tiled_jc() { return g_tiled_jc; }

constexpr size_t kMaxLds = 160 * 1024;

bool tiled_supported(const Grid& g, size_t elem_bytes) {
  if (g.dim != 3) return false;
  const TiledArgs a = tiled_args(g, 1, 1, g_tiled_jc);
  const size_t bytes = (size_t)12 * a.u_span * elem_bytes + (size_t)3 * a.e_span * elem_bytes;
  return tiled_shape_ok(a) && bytes <= kMaxLds;
}

template <typename T, int MODE>
static void launch_tiled(dim3 grid, size_t lds, hipStream_t s, const TiledArgs& A, const T* K0,
                         const T* E, const uint8_t* cons, const T* x, const T* b, const T* Dinv,
                         const double* omega, double wscale, const T* xadd, T* y) {
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&op0_tiled_kernel<T, MODE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
    attr_set = true;
  }
  op0_tiled_kernel<T, MODE><<<grid, kStrip, lds, s>>>(A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y);
}

template <typename T>
int op0_tiled(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, const T* K0, const T* E,
              const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
              double wscale, const T* xadd, T* y, hipStream_t s) {
  const TiledArgs A = tiled_args(g, e_sx, e_sy, g_tiled_jc);
  const size_t lds = tiled_lds_bytes<T>(A);
  ET_REQUIRE(tiled_shape_ok(A) && lds <= kMaxLds, "tiled operator: grid too wide in z for the LDS strip");
  const int64_t plane = (int64_t)g.nx() * g.nz();
  dim3 grid((unsigned)((plane + kStrip - 1) / kStrip), (unsigned)((g.ny() + A.jc - 1) / A.jc));
  switch (mode) {
    case TM_Y: launch_tiled<T, TM_Y>(grid, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_RES: launch_tiled<T, TM_RES>(grid, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_JAC: launch_tiled<T, TM_JAC>(grid, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    default: launch_tiled<T, TM_JAC_ADD>(grid, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

template int op0_tiled<double>(int, const Grid&, int64_t, int64_t, const double*, const double*,
                               const uint8_t*, const double*, const double*, const double*,
                               const double*, double, const double*, double*, hipStream_t);
template int op0_tiled<float>(int, const Grid&, int64_t, int64_t, const float*, const float*,
                              const uint8_t*, const float*, const float*, const float*,
                              const double*, double, const float*, float*, hipStream_t);

}  // namespace et
