// The opaque handles of the C ABI (include/mdc.h) over the engine state.
#pragma once
#include "engine.h"
#include "../../include/mdc.h"

struct mdc_ctx {
    mdc::Ctx c;
};
struct mdc_pot {
    mdc::Pot p;
};
struct mdc_sys {
    mdc::Batch b;
};

namespace mdc {
// 208-byte record (mdcraft/data/atom.h:21-38)
struct AtomRec {
    double r[3], v[3], f[3], V[9];
    double m, rcut, rns, Ep, Em, nc;
    uint16_t bc, kind;
    uint32_t uid;
    uint16_t tag, pad[3];
};
static_assert(sizeof(AtomRec) == 208, "atom record must be 208 bytes");

// Records the message as the calling thread's last error (and the
// context's, when given) and returns code.
int report_error(Ctx* c, int code, const std::string& msg);
}  // namespace mdc
