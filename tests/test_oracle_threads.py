"""The oracle's z-slab threads (oracle_set_threads, the analogue of the
reference's MPI z partition) give the sequential oracle's results bit for
bit: every cell and face is computed by the same operations, the scan-order
output, excitation and face numbers start each plane from precomputed
counts."""
import numpy as np
import pytest

import problems
from ggi_tlm_amd.problem import build_tables

DECKS = {
    "pml_sphere": problems.pml_sphere,
    "special_mix": problems.special_mix,
    "face_outputs": problems.face_outputs,
    "dispersive_blocks": problems.dispersive_blocks,
    "pml_material": lambda: problems.pml_material(steps=120),
    "huygens_face_outputs": problems.huygens_face_outputs,
}


@pytest.mark.parametrize("name", sorted(DECKS))
def test_threads_bitwise(oracle_lib, name):
    from oracle import oracle

    t = build_tables(DECKS[name]())
    steps = min(t.n_timesteps, 120)
    res = []
    for threads in (1, 3, 8):
        h = oracle.handle(t, output_capacity=steps, threads=threads)
        if name == "pml_sphere":
            h.set_small_to_zero(7, 1e-9)
        h.step(1, steps)
        r = [h.read_cell_outputs()[0], h.download_links()]
        if t.n_output_faces:
            r.append(h.read_face_outputs()[0])
        if len(t.dft_frequencies) and len(t.output_rows):
            r.append(h.read_dft())
        h.close()
        res.append(r)
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert a.shape == b.shape and np.array_equal(a, b)


def test_set_threads_rejects_zero(oracle_lib):
    from oracle import oracle

    t = build_tables(problems.free_cube(6, 4))
    with pytest.raises(ValueError):
        oracle.handle(t, threads=0)
    assert oracle.max_threads() >= 1
