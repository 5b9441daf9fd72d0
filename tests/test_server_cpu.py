"""Request-loop admission on the host: footprint model and over-budget rejection
(no device needed: requests are validated and sized before they are queued)."""
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
