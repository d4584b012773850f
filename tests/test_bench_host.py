"""Host logic of bench.py that needs no GPU: the sizing of the CPU-baseline
sample."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402


def _fake_oracle(per_tuple, setup, noise):
    """A stand-in for the timed oracle run: a one-time setup whose duration
    wobbles by `noise(count)` seconds, plus per_tuple seconds per tuple and
    thread; records the tuple counts it was asked for."""
    calls = []

    def run(O, no, nv, t, first, count, threads):
        calls.append(count)
        return setup + noise(count) + count * per_tuple / threads, 0.0

    return run, calls


def test_probe_grows_past_setup_noise(monkeypatch):
    # a 13 ms tuple behind a 6 s setup that reads 0.3 s long on the empty
    # run and 0.2 s short on small probes: a 4-tuple probe then shows a
    # negative loop time, which once sized the sample to 400,000 tuples
    run, calls = _fake_oracle(
        0.013, 6.0, lambda n: 0.3 if n == 0 else (-0.2 if n < 100 else 0.0))
    monkeypatch.setattr(bench, "_timed_oracle", run)
    count, dt, setup = bench._sized_sample(None, 40, 400, None, 1, 10.0)
    assert 400 <= count <= 2000          # about 10 s of 13 ms tuples
    assert max(calls) <= count            # no probe beyond the sample
    assert dt > 0 and setup > 0


def test_sample_is_a_multiple_of_the_threads(monkeypatch):
    run, _ = _fake_oracle(0.02, 1.0, lambda n: 0.0)
    monkeypatch.setattr(bench, "_timed_oracle", run)
    count, _, _ = bench._sized_sample(None, 40, 400, None, 16, 5.0)
    assert count % 16 == 0
    assert 0.5 * 5.0 <= count * 0.02 / 16 <= 2.0 * 5.0


def test_sample_never_exceeds_the_space(monkeypatch):
    # tiny space: the probe reaches all of it and stops there
    run, calls = _fake_oracle(1e-6, 0.5, lambda n: 0.0)
    monkeypatch.setattr(bench, "_timed_oracle", run)
    count, _, _ = bench._sized_sample(None, 2, 6, None, 1, 10.0)
    assert count <= bench.n_tuples(6)
    assert max(calls) <= bench.n_tuples(6)


def test_more_than_one_gpu_is_refused(capsys, monkeypatch):
    # one GPU only: --gpus above 1, or a launcher's WORLD_SIZE above 1, ends
    # before any device work with a message that says so
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err
    monkeypatch.setenv("WORLD_SIZE", "8")
    assert bench.main([]) == 2
