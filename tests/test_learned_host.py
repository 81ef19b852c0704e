"""Host logic of the learned initialiser: overlap pair list and its CSR."""
import numpy as np

from ptychography_amd.learned import overlap_pairs, pair_csr


def test_overlap_pairs_brute_force():
    rng = np.random.default_rng(0)
    off = rng.integers(0, 300, size=(1500, 2))   # > one 1024-row block
    pairs = overlap_pairs(off, 64, 48, frac=0.75)
    ref = [(i, j) for i in range(len(off)) for j in range(i + 1, len(off))
           if abs(off[j, 0] - off[i, 0]) < 48 and abs(off[j, 1] - off[i, 1]) < 36]
    assert [tuple(p) for p in pairs] == ref


def test_pair_csr_is_the_hermitian_pattern():
    rng = np.random.default_rng(1)
    off = rng.integers(0, 120, size=(60, 2))
    pairs = overlap_pairs(off, 64, 64)
    rowptr, ent = pair_csr(pairs, len(off))
    assert rowptr[0] == 0 and rowptr[-1] == 2 * len(pairs)
    seen = set()
    for i in range(len(off)):
        cols = ent[rowptr[i]:rowptr[i + 1], 0]
        assert np.all(np.diff(cols) > 0)
        for j, qs in ent[rowptr[i]:rowptr[i + 1]]:
            q = qs if qs >= 0 else -qs - 1
            assert tuple(pairs[q]) == ((i, j) if qs >= 0 else (j, i))
            seen.add((i, j))
    assert len(seen) == 2 * len(pairs)


def _dihedral(a, code):
    """The index map of the training-batch kernel: out[i, j] = a[T(i, j)] with
    rows i -> (-i) mod M (bit 0), columns (bit 1), then a transpose (bit 2)."""
    M = a.shape[0]
    i, j = np.meshgrid(np.arange(M), np.arange(M), indexing="ij")
    si = (M - i) % M if code & 1 else i
    sj = (M - j) % M if code & 2 else j
    if code & 4:
        si, sj = sj, si
    return a[si, sj]


def test_dihedral_maps_are_exact_symmetries_of_the_training_forward_model():
    """For the training probe (disc pupil + defocus), reflecting / transposing
    the object patch maps the centred diffraction pattern the same way."""
    from ptychography_amd import sim
    rng = np.random.default_rng(3)
    P = sim.make_probe(64, 16, 0.02, 0.0, rng)
    O = sim.make_object((64, 64), rng, (0.3, 1.0), 4.0)

    def intensity(o):
        return np.fft.ifftshift(np.abs(np.fft.fft2(P * o)) ** 2)

    I0 = intensity(O)
    for code in range(8):
        assert np.allclose(intensity(_dihedral(O, code)), _dihedral(I0, code), rtol=1e-9,
                           atol=1e-9 * I0.max())
    # a probe without the symmetry (random Zernike phase) breaks it
    Pz = sim.make_probe(64, 16, 0.02, 0.6, np.random.default_rng(4))
    Iz = np.fft.ifftshift(np.abs(np.fft.fft2(Pz * O)) ** 2)
    Iz1 = np.fft.ifftshift(np.abs(np.fft.fft2(Pz * _dihedral(O, 1))) ** 2)
    assert not np.allclose(Iz1, _dihedral(Iz, 1), rtol=1e-3)
