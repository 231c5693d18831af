"""The oracle's stratified sampler (oracle.c strat2), pinned by what the
construction guarantees rather than by its own formula: shuffled,
Owen-scrambled Sobol points (Burley 2020) make every aligned power-of-two
block of a pixel's samples a (0, m, 2)-net in base 2, i.e. each of the 2^m
elementary intervals of every shape 2^-a x 2^-(m-a) holds exactly one point.
A wrong Sobol generator (dimension 1 equal to dimension 0, a shifted column
rule), a scramble that is not nested, or an index shuffle that breaks
aligned blocks all fail it. The integrator side is checked by convergence
to the independent sampler's mean and by its lower error at equal spp."""
import numpy as np
import pytest

from conftest import aqua, oracle, small_camera


def _cells(pts, a, b):
    x = np.floor(pts[:, 0].astype(np.float64) * (1 << a)).astype(np.int64)
    y = np.floor(pts[:, 1].astype(np.float64) * (1 << b)).astype(np.int64)
    return x * (1 << b) + y


@pytest.mark.parametrize("seed,frame,pixel", [(0, 0, 0), (7, 3, 12345), (0xFFFFFFFF, 239, 2073599)])
@pytest.mark.parametrize("pattern", [0, 1, 2, 3])
def test_aligned_power_of_two_blocks_are_nets(seed, frame, pixel, pattern):
    pts = oracle.strat2(seed, frame, pixel, pattern, 512)
    assert pts.min() >= 0.0 and pts.max() < 1.0
    for m in range(10):
        n = 1 << m
        for k in range(512 // n):
            blk = pts[k * n:(k + 1) * n]
            for a in range(m + 1):
                c = _cells(blk, a, m - a)
                assert len(np.unique(c)) == n, (m, k, a)


def test_blocks_far_into_the_sequence_are_nets():
    # a later pass of a frame starts at its sample index: blocks there are
    # nets too (the index shuffle maps aligned blocks onto aligned blocks)
    pts = oracle.strat2(3, 17, 999, 2, 256, first=(1 << 19) + 256)
    for m in (4, 8):
        n = 1 << m
        for k in range(256 // n):
            for a in range(m + 1):
                assert len(np.unique(_cells(pts[k * n:(k + 1) * n], a, m - a))) == n


def test_patterns_are_decorrelated_and_uniform():
    a = oracle.strat2(0, 0, 5, 1, 64)
    for other in (oracle.strat2(0, 0, 6, 1, 64), oracle.strat2(0, 1, 5, 1, 64), oracle.strat2(0, 0, 5, 2, 64),
                  oracle.strat2(1, 0, 5, 1, 64)):
        assert not np.array_equal(a, other)
        # not the same point set in another order either
        assert not np.array_equal(np.sort(a[:, 0]), np.sort(other[:, 0]))
    # sample 0 over many pixels: uniform (chi^2 over 16 x 16 cells, 99.9 % bound)
    first = np.array([oracle.strat2(0, 0, p, 0, 1)[0] for p in range(4096)])
    h, _, _ = np.histogram2d(first[:, 0], first[:, 1], bins=16, range=[[0, 1], [0, 1]])
    chi2 = float(((h - 16.0) ** 2 / 16.0).sum())
    assert chi2 < 346.0, chi2  # chi^2_255 at p = 0.999


def test_stratified_render_converges_to_the_same_mean_with_lower_error(box_room, box_room_oracle):
    cam = aqua.camera_to_array(small_camera(box_room.camera(), 40, 40))
    ref = box_room_oracle.render64(cam, 4096, 1, seed=11, frame=900)
    strat = box_room_oracle.render64(cam, 1024, 1, seed=11, frame=1, sampler=1)
    # same expectation: image means agree to the noise of the two renders
    assert abs(strat.mean() / ref.mean() - 1.0) < 0.01

    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref * ref + 1e-2)))

    e_ind = np.mean([relmse(box_room_oracle.render(cam, 16, 1, seed=11, frame=f)[0]) for f in range(4)])
    e_str = np.mean([relmse(box_room_oracle.render(cam, 16, 1, seed=11, frame=f, sampler=1)[0]) for f in range(4)])
    # direct light at 16 spp: the stratified patterns measured 2-3x lower
    assert e_str < 0.6 * e_ind, (e_str, e_ind)


def test_integrator_json_sampler_key(tmp_path):
    p = tmp_path / "integrator.json"
    p.write_text('{"type": "pt", "spp": 4, "sampler": "stratified"}')
    assert aqua.integrator_load(str(p)).sampler == aqua.SAMPLER_STRATIFIED
    p.write_text('{"type": "pt", "spp": 4}')
    assert aqua.integrator_load(str(p)).sampler == aqua.SAMPLER_INDEPENDENT
    p.write_text('{"type": "pt", "spp": 4, "sampler": "sobol"}')
    with pytest.raises(aqua.AquaError):
        aqua.integrator_load(str(p))
