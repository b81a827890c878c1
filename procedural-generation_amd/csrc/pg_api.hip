// C-ABI plumbing: the per-thread error message, the ABI version, and the
// host-side entry points of the library's scalar geometry (the same
// double-double tan/atan and cube-sphere code the kernels run).
#include <stdarg.h>
#include <stdio.h>

#include "pg_common.h"
#include "pg_erode.h"
#include "pg_internal.h"

namespace pg {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_error[0] = '\0'; }

}  // namespace pg

using namespace pg;

extern "C" const char* pg_last_error(void) { return g_error; }

extern "C" int pg_abi_version(void) { return 9; }

extern "C" int pg_geom_to_sphere(int face, double u, double v, double* xyz3) {
    clear_error();
    PG_REQUIRE(face >= 0 && face < 6 && xyz3, "pg_geom_to_sphere: bad arguments");
    Vec3 p = to_sphere(face, u, v);
    xyz3[0] = p.x;
    xyz3[1] = p.y;
    xyz3[2] = p.z;
    return 0;
}

extern "C" int pg_geom_from_sphere(double x, double y, double z, int* face, double* uv2) {
    clear_error();
    PG_REQUIRE(face && uv2, "pg_geom_from_sphere: bad arguments");
    FaceUV q = from_sphere(x, y, z);
    *face = q.face;
    uv2[0] = q.u;
    uv2[1] = q.v;
    return 0;
}

extern "C" int pg_geom_jacobian(int face, double u, double v, double* j6) {
    clear_error();
    PG_REQUIRE(face >= 0 && face < 6 && j6, "pg_geom_jacobian: bad arguments");
    jacobian(face, u, v, j6);
    return 0;
}

extern "C" int pg_geom_transfer_velocity(int face, double u, double v, double a, double b, int* face_out, double* uvab4) {
    clear_error();
    PG_REQUIRE(face >= 0 && face < 6 && face_out && uvab4, "pg_geom_transfer_velocity: bad arguments");
    Transfer t = transfer_velocity(face, u, v, a, b);
    *face_out = t.face;
    uvab4[0] = t.u;
    uvab4[1] = t.v;
    uvab4[2] = t.a;
    uvab4[3] = t.b;
    return 0;
}

extern "C" int pg_geom_transfer_velocity_f32(int face, float u, float v, float a, float b, int* face_out, float* uvab4) {
    clear_error();
    PG_REQUIRE(face >= 0 && face < 6 && face_out && uvab4, "pg_geom_transfer_velocity_f32: bad arguments");
    const TransferF t = transfer_velocity_f(face, u, v, a, b);
    *face_out = t.face;
    uvab4[0] = t.u;
    uvab4[1] = t.v;
    uvab4[2] = t.a;
    uvab4[3] = t.b;
    return 0;
}

extern "C" double pg_geom_tan(double x) { return pg_tan(x); }
extern "C" double pg_geom_atan(double x) { return pg_atan(x); }
extern "C" void pg_geom_sincos(double x, double* s, double* c) { pg_sincos(x, s, c); }

namespace pg {
namespace {
// Device evaluation of the scalar geometry over arrays (the kernels' own
// code path, for checking it on the GPU against golden vectors).
__global__ void geom_eval_kernel(int op, int64_t n, const double* in, double* out, int* iout) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        if (op == 0) {
            out[k] = pg_tan(in[k]);
        } else if (op == 1) {
            out[k] = pg_atan(in[k]);
        } else if (op == 2) {
            Vec3 p = to_sphere((int)in[3 * k], in[3 * k + 1], in[3 * k + 2]);
            out[3 * k] = p.x;
            out[3 * k + 1] = p.y;
            out[3 * k + 2] = p.z;
        } else if (op == 3) {
            FaceUV q = from_sphere(in[3 * k], in[3 * k + 1], in[3 * k + 2]);
            iout[k] = q.face;
            out[2 * k] = q.u;
            out[2 * k + 1] = q.v;
        } else if (op == 4) {
            Transfer t = transfer_velocity((int)in[5 * k], in[5 * k + 1], in[5 * k + 2], in[5 * k + 3], in[5 * k + 4]);
            iout[k] = t.face;
            out[4 * k] = t.u;
            out[4 * k + 1] = t.v;
            out[4 * k + 2] = t.a;
            out[4 * k + 3] = t.b;
        } else if (op == 5) {
            pg_sincos(in[k], &out[2 * k], &out[2 * k + 1]);
        } else if (op == 6) {
            const TransferF t = transfer_velocity_f((int)in[5 * k], (float)in[5 * k + 1], (float)in[5 * k + 2],
                                                   (float)in[5 * k + 3], (float)in[5 * k + 4]);
            iout[k] = t.face;
            out[4 * k] = t.u;
            out[4 * k + 1] = t.v;
            out[4 * k + 2] = t.a;
            out[4 * k + 3] = t.b;
        }
    }
}
}  // namespace
}  // namespace pg

extern "C" int pg_geom_eval(int op, int64_t n, const double* in, double* out, int* iout, pg_stream stream) {
    clear_error();
    PG_REQUIRE(op >= 0 && op <= 6 && n >= 0, "pg_geom_eval: bad op %d or n", op);
    if (n == 0) return 0;
    PG_REQUIRE(in && out && (iout || (op != 3 && op != 4 && op != 6)), "pg_geom_eval: NULL array");
    int blocks = blocks_for(n, 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(geom_eval_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, op, n, in, out, iout);
    PG_LAUNCHED("geom_eval_kernel");
    return 0;
}
