// compgt_mfma.hip -- int8 MFMA form of the ground-truth scan (placeholder
// until the MFMA tile kernel lands).
#include "internal.h"

namespace nbpg {

int compgt_mfma(int dtype, const uint8_t *base, uint64_t n, const uint8_t *queries, uint64_t nq,
                uint32_t dims, uint32_t k, uint64_t *partial, uint32_t slices, uint64_t slice_rows,
                hipStream_t stream)
{
    (void)dtype; (void)base; (void)n; (void)queries; (void)nq; (void)dims; (void)k;
    (void)partial; (void)slices; (void)slice_rows; (void)stream;
    return set_error(NBPG_EUNSUP, "MFMA ground truth not built yet");
}

} // namespace nbpg
