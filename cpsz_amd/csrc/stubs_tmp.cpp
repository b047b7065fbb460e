#include "codec3d.hpp"
#include "cp.hpp"
namespace cpsz {
void critical_points_2d(Ctx&, const float*, const float*, size_t, size_t, int64_t*, int32_t*, size_t, size_t*) { fail(kErrUnsupported, "not built yet"); }
void critical_points_3d(Ctx&, const float*, const float*, const float*, size_t, size_t, size_t, int64_t*, int32_t*, size_t, size_t*) { fail(kErrUnsupported, "not built yet"); }
}
