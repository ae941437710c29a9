// fmm.hip -- octree fast multipole method for the demag potential (placeholder until the
// hierarchical solver lands; FG_DEMAG_FMM is refused at fg_create).
#include "ctx.h"

struct FmmPlan {
    int unused;
};

namespace fg {

int fmm_create(Ctx &c)
{
    c.err = "FMM demag is not available in this build";
    return FG_ERR_ARG;
}

void fmm_destroy(Ctx &c)
{
    delete c.fmm;
    c.fmm = nullptr;
}

void launch_fmm(Ctx &) {}

}  // namespace fg
