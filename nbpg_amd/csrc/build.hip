// build.hip -- build_vamana (nsw.h:547-663) on the device (in progress).
#include "internal.h"

using namespace nbpg;

extern "C" int nbpg_build(nbpg_index *idx, const nbpg_build_params *bp, uint64_t *out_isolated)
{
    (void)idx; (void)bp; (void)out_isolated;
    return set_error(NBPG_EUNSUP, "device build not available yet");
}
