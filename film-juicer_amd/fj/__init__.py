"""Host-side Python mirror of the Film-Juicer native render boundary (MI355X).

`fj.abi` mirrors the C-ABI records, `fj.prepared` builds prepared-host-data
records from prepared-product case files, `fj.runtime` loads the HIP library
and drives `fj_cuda_*`, `fj.synth` generates seeded synthetic frames.
"""
