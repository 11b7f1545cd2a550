#include "dna_internal.h"
#include "dna_ctx.h"
extern "C" int dna_score(dna_ctx *, const dna_dna_target *, const float *, uint64_t, uint32_t *, int, dna_score_out *) { dna::set_error("not built"); return DNA_ERR_ARG; }
extern "C" int dna_hash_cells(dna_ctx *, const float *, uint64_t, float, uint32_t, uint32_t *, uint32_t *, uint32_t *) { dna::set_error("not built"); return DNA_ERR_ARG; }
