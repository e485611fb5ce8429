"""The guided stepper's guess policy (lcode3d_amd.lcode.GuessBackoff), on
the CPU: after a rejected guess the plain step runs for a waiting time that
doubles with every further rejection up to a cap, and an accepted guess
restores guessing on every slice."""

from lcode3d_amd.lcode import GuessBackoff


def _run(policy, outcomes, slices):
    """Guided (G) or plain (.) per slice; outcomes feed the guided ones."""
    out, it = [], iter(outcomes)
    for _ in range(slices):
        if policy.want_guess():
            out.append("G")
            policy.observe(next(it))
        else:
            out.append(".")
    return "".join(out)


def test_backoff_doubles_up_to_the_cap():
    p = GuessBackoff(backoff=2, max_backoff=8)
    assert _run(p, [False] * 5, 1 + 2 + 1 + 4 + 1 + 8 + 1 + 8 + 1) == \
        "G..G....G........G........G"


def test_acceptance_restores_every_slice_and_the_initial_wait():
    p = GuessBackoff(backoff=2, max_backoff=8)
    seq = _run(p, [False, False, True, True, True, False], 14)
    # reject: wait 2; reject: wait 4; then accepted three times in a row;
    # the next rejection waits 2 again
    assert seq == "G..G....GGGG..", seq


def test_zero_backoff_guesses_every_slice():
    p = GuessBackoff(backoff=0, max_backoff=64)
    assert _run(p, [False] * 10, 10) == "G" * 10
