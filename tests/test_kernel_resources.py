"""Register and scratch budgets of the hot kernels, read from the gfx950 assembly metadata.

The two-step acoustic kernels and the elastic step kernels are written for three wavefronts per
SIMD: gfx950 gives a SIMD 512 VGPRs per lane, so a kernel of at most 168 VGPRs (allocated in
blocks of 8) keeps three waves resident, and one VGPR more leaves two.  Their loops are latency
bound, so losing that wave, or spilling registers to scratch memory, costs a large part of their
throughput without changing any result.  The benchmark workloads run exactly the variants listed
here (acoustic: 4 shots per forward wavefront, 2 per adjoint wavefront; elastic: the
non-anisotropic instantiations, which the isotropic and VTI models use).  Host-only: hipcc
cross-compiles the device code, no GPU is needed.
"""
import os
import re
import shutil

import pytest

from adfwi_amd import build

MAX_VGPRS_3_WAVES = 168

# kernel-name fragments (mangled) of the variants the benchmarks run, per source file
BUDGETED = {
    "acoustic.hip": [
        "ac_fwd_pairILi4ELi1E",          # records and energy (observed data, gradient head)
        "ac_fwd_pairILi4ELi2E",          # recompute: stencil images, branch-free stores
        "ac_fwd_pairILi4ELi10E",         # recompute, stores in branches (fields of 1 GiB and more)
        "ac_fwd_pairILi4ELi3E",          # records, energy and images (resident segments)
        "ac_adj_pairILi2ELb0ELb0ELb1E",  # vp gradient from pressure records, branch-free stores
        "ac_adj_pairILi2ELb0ELb0ELb0E",  # the same with stores in branches
        "ac_adj_pairILi2ELb0ELb1ELb1E",  # vp and wavelet gradients from all records
        "ac_adj_pairILi2ELb1ELb0ELb0E",  # with the rho gradient
    ],
    "elastic.hip": [
        "el_fwd_stepILb0ELb1E", "el_fwd_stepILb0ELb0E",
        "el_adj_stepILb0ELb1E", "el_adj_stepILb0ELb0E",
        "el3_fwd_stepILb0E", "el3_adj_stepILb0E",
    ],
}


def _kernels(asm_text):
    md = asm_text[asm_text.index("amdhsa.kernels:"):]
    out = {}
    for entry in re.split(r"\n  - ", md)[1:]:
        name = re.search(r"\.name:\s+(\S+)", entry)
        if not name:
            continue
        field = lambda key: int(re.search(r"\." + key + r":\s+(\d+)", entry).group(1))
        out[name.group(1)] = {k: field(k) for k in
                              ("vgpr_count", "vgpr_spill_count", "private_segment_fixed_size")}
    return out


@pytest.mark.skipif(shutil.which(build.hipcc()) is None and not os.path.exists(build.hipcc()),
                    reason="hipcc not available")
@pytest.mark.parametrize("source", sorted(BUDGETED))
def test_hot_kernels_keep_three_waves_per_simd_without_scratch(source, tmp_path):
    asm = open(build.device_asm(source, str(tmp_path / (source + ".s")))).read()
    kernels = _kernels(asm)
    for frag in BUDGETED[source]:
        hits = [n for n in kernels if frag in n]
        assert len(hits) == 1, (frag, hits)
        k = kernels[hits[0]]
        assert k["vgpr_count"] <= MAX_VGPRS_3_WAVES, (frag, k)
        assert k["vgpr_spill_count"] == 0 and k["private_segment_fixed_size"] == 0, (frag, k)
