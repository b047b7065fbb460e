#include "codec3d.hpp"
#include "cp.hpp"
namespace cpsz {
void compress_3d_device(Ctx&, const float*, const float*, const float*, size_t, size_t, size_t, double, DeviceStream&) { fail(kErrUnsupported, "3D not built yet"); }
void decompress_3d_device(Ctx&, const uint8_t*, const uint8_t*, size_t, size_t, size_t, size_t, float*, float*, float*) { fail(kErrUnsupported, "3D not built yet"); }
void critical_points_2d(Ctx&, const float*, const float*, size_t, size_t, int64_t*, int32_t*, size_t, size_t*) { fail(kErrUnsupported, "not built yet"); }
void critical_points_3d(Ctx&, const float*, const float*, const float*, size_t, size_t, size_t, int64_t*, int32_t*, size_t, size_t*) { fail(kErrUnsupported, "not built yet"); }
const LibmTables& libm_tables() { static LibmTables t; return t; }
}
