// The chemical mechanism of this build.  BRAMS compiles one mechanism in
// (its chem1_list and SPACK kernels, src/brams/ccatt/<MECH>/); so does this
// library: BRAMS_MECH_GEN names the generated device functions and
// BRAMS_MECH_SLAB the slab-layout variants, both from
// brams_amd/codegen/gen_mech.py.  RELACS_TUV by default.
#pragma once
#ifndef BRAMS_MECH_GEN
#define BRAMS_MECH_GEN "generated/relacs_tuv_gen.h"
#define BRAMS_MECH_SLAB "generated/relacs_tuv_slab.h"
#endif
#include BRAMS_MECH_GEN
