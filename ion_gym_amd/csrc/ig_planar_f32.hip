// planar flight kernels, IG_PREC_F32 instantiations (one translation unit
// per precision so the kernels compile in parallel).
#include "ig_planar_kernel.h"

namespace ig {
template int launch_planar_prec<IG_PREC_F32>(const FieldsK &, const FlightK &, const EnsK &,
                                            int64_t, int, void *);
} // namespace ig
