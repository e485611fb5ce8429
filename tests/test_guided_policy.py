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


def test_alternating_extrapolation_coefficients():
    """2 F(n-1) - 2 F(n-3) + F(n-4) is exact for fields quadratic in xi and
    carries an alternating part e (-1)^k on as e (-1)^n, where the
    quadratic extrapolation 3 F(n-1) - 3 F(n-2) + F(n-3) gives -7 e (-1)^n."""
    import types

    import numpy as np

    from lcode3d_amd.lcode import FieldExtrapolator

    def state(v):
        return types.SimpleNamespace(Ex=np.float64(v))

    def apply(terms):
        return sum(c * s.Ex for c, s in terms)

    n = 10
    for quad in (lambda k: 1.0, lambda k: 0.5 * k, lambda k: 0.25 * k * k - k + 3):
        for alt in (False, True):
            e = FieldExtrapolator(order=2, alternating=alt)
            for k in range(n):
                e.push(state(quad(k)))
            assert abs(apply(e.guess()) - quad(n)) < 1e-9
    e_alt, e_quad = (FieldExtrapolator(order=2, alternating=a) for a in (True, False))
    for k in range(n):
        for e in (e_alt, e_quad):
            e.push(state((-1.0) ** k))
    assert abs(apply(e_alt.guess()) - (-1.0) ** n) < 1e-12
    assert abs(apply(e_quad.guess()) + 7 * (-1.0) ** n) < 1e-12
    # before four slices are known the alternating form falls back
    e = FieldExtrapolator(order=2, alternating=True)
    for k in range(3):
        e.push(state(k * k))
    assert len(e.guess()) == 3 and abs(apply(e.guess()) - 9) < 1e-12
