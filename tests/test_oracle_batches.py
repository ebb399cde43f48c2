"""Pins the oracle's restatement of estimateNumBatches and of the reference's
device-memory accounting (interaction_elph.cpp:458-490, 703-707) with
hand-evaluated cases."""
from oracle import oracle as O


def test_accounted_bytes_hand_evaluated():
    # nw=2, nm=3, nRp=5, nRe=7: 16 * 420 + 8 * (5 + 15 + 21 + 7) = 6720 + 384
    assert O.accounted_bytes(2, 3, 5, 7) == 7104
    # after cacheElPh with nb1 = 1: elPhCached holds 5 * 3 * 1 * 2 = 30 elements
    assert O.accounted_bytes(2, 3, 5, 7, nb1_cached=1) == 7104 + 480


def test_estimate_num_batches_hand_evaluated():
    # per q: evs 16*(4+9)=208, polar 16*3*1*2=96, max(phase+g3=80+192,
    # g3+g4=384, g4+gFinal=384, gFinal+coupling=192+96) = 384: 688 bytes
    assert O.estimate_num_batches(2, 3, 5, 10, 1, 1000.0) == 7     # ceil(6880 / 1000)
    assert O.estimate_num_batches(2, 3, 5, 10, 1, 700.0) == 10     # ceil(9.83)
    assert O.estimate_num_batches(2, 3, 5, 10, 1, 688.0) == 10     # exactly one q per batch
    assert O.estimate_num_batches(2, 3, 5, 10, 1, 6880.0) == 1
    assert O.estimate_num_batches(2, 3, 5, 10, 1, 600.0) is None   # 600 < 688: one q does not fit
    assert O.estimate_num_batches(2, 3, 5, 0, 1, 600.0) == 0       # no q: no batch, no error
    # nb1 = 2 doubles the band-linear terms: per q 208 + 192 + 768 = 1168
    assert O.estimate_num_batches(2, 3, 5, 3, 2, 1168.0) == 3
