"""Request-loop admission on the host: footprint model and over-budget rejection
(no device needed: requests are validated and sized before they are queued)."""
import numpy as np
import pytest

from helpers import sphere_case


def test_footprint_grows_with_nmax_and_precision():
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    f1 = native.footprint(c, e, p, 1)
    f3 = native.footprint(c, e, p, 3)
    fm = native.footprint(c, e, p, 3, precision=native.PREC_MIXED)
    assert 0 < f1 < f3 < fm


def test_footprint_rejects_bad_mesh():
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    e = e.copy()
    e[0, [0, 1]] = e[0, [1, 0]]
    with pytest.raises(native.ScadynError) as ex:
        native.footprint(c, e, p, 2)
    assert ex.value.code == native.ERR_ORIENTATION


def test_over_budget_request_is_rejected_not_queued():
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    sv = native.Server(hbm_budget=1 << 20)          # 1 MiB: smaller than any request
    with pytest.raises(native.ScadynError) as ex:
        sv.submit(c, e, p, k, 2)
    assert ex.value.code == native.ERR_MEMORY
    sv.close()


def test_host_only_loop_admits_and_retires_concurrently():
    """Submit and wait from several threads against the loop's host-only
    stand-in solve: every request is admitted within the budget, retired once,
    and its waiter gets its own record (the stand-in's iteration counts carry
    the request id).  Waiters free their requests while the worker keeps
    stepping others, the pattern that a publish-before-retire race breaks."""
    import threading
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    fp = native.footprint(c, e, p, 1)
    # room for two requests at a time
    sv = native.Server(max_active=8, hbm_budget=int(2.5 * fp), host_only=True)
    errors = []
    lock = threading.Lock()

    def client(nreq):
        try:
            for _ in range(nreq):
                with lock:
                    rid = sv.submit(c, e, p, k, 1)
                T, it, times = sv.wait(rid)
                assert (it == rid).all(), (rid, it)
                assert times[0] <= times[1] <= times[2] <= times[3]
                assert all(np.linalg.norm(t) == 0 for t in T)
        except Exception as ex:  # pragma: no cover - reported below
            errors.append(repr(ex))

    threads = [threading.Thread(target=client, args=(6,)) for _ in range(6)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    sv.close()
    assert not errors, errors
    assert not any(t.is_alive() for t in threads)


def test_host_only_cancel_queued_and_running():
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    fp = native.footprint(c, e, p, 1)
    sv = native.Server(max_active=1, hbm_budget=int(1.5 * fp), host_only=True)
    a = sv.submit(c, e, p, k, 1)
    b = sv.submit(c, e, p, k, 1)
    sv.cancel(b)                                   # queued behind a: leaves the queue
    with pytest.raises(native.ScadynError, match='cancelled'):
        sv.wait(b)
    T, it, _ = sv.wait(a)                          # a is unaffected
    assert (it == a).all()
    d = sv.submit(c, e, p, k, 1)
    sv.cancel(d)
    with pytest.raises(native.ScadynError, match='cancelled'):
        sv.wait(d)
    sv.close()


def test_footprint_of_the_mixed_basis_storage():
    """The footprint sizes the Krylov basis the solve will hold: binary16 when
    asked for, complex64 under the automatic choice below 2^20 rows (this
    mesh), and the complex128 basis of the FP64 solve whatever `basis` says."""
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    mix = dict(precision=native.PREC_MIXED, batch=16)
    f_auto = native.footprint(c, e, p, 3, **mix)
    f_c64 = native.footprint(c, e, p, 3, basis=native.BASIS_C64, **mix)
    f_f16 = native.footprint(c, e, p, 3, basis=native.BASIS_F16, **mix)
    assert f_auto == f_c64
    # per slot: 51 basis vectors of 4 bytes less per element, one complex64
    # vector more (the matvec input converted from the basis)
    n = 3 * e.shape[0]
    assert f_c64 - f_f16 == 16 * (51 * n * 4 - n * 8)
    assert native.footprint(c, e, p, 3, basis=native.BASIS_F16) == native.footprint(c, e, p, 3)


def test_invalid_options_fail_on_the_host():
    """Option values the solver does not know are rejected by the footprint
    query and at submission, before a request reaches the device."""
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    for bad in (dict(basis=7), dict(gather_kernel=2), dict(mgs_block=3), dict(precision=5)):
        with pytest.raises(native.ScadynError) as ex:
            native.footprint(c, e, p, 2, **bad)
        assert ex.value.code == native.ERR_ARG
    sv = native.Server(hbm_budget=1 << 40)
    with pytest.raises(native.ScadynError) as ex:
        sv.submit(c, e, p, k, 2, basis=7)
    assert ex.value.code == native.ERR_ARG
    sv.close()
