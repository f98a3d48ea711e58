// placeholder: BPNN kernels are added in the next change
#include "engine.h"
namespace mdc {
void bpnn_compute(Batch&, const Pot&, int, int, int, double, int) { throw Error(E_STATE, "BPNN not built yet"); }
Pot* bpnn_create(Ctx*, const char*) { throw Error(E_STATE, "BPNN not built yet"); }
void bpnn_destroy(Pot* p) { delete p; }
}
