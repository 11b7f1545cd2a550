#include "dna_internal.h"
#include "dna_ctx.h"
extern "C" int dna_run_irt(dna_ctx *, const float *, uint64_t, uint32_t, float, const double *, int, uint64_t, dna_chem_out *) { dna::set_error("not built"); return DNA_ERR_ARG; }
extern "C" int dna_run_irt_batch(dna_ctx *, const float *, const uint64_t *, uint32_t, const uint32_t *, const float *, const double *, int, uint64_t, dna_chem_out *) { dna::set_error("not built"); return DNA_ERR_ARG; }
