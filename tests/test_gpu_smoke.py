"""The package's smoke entry (one small device-resident run and one drop-in
step against the FP64 oracle) on the GPU."""
import pytest

pytestmark = pytest.mark.gpu


def test_smoke_entry_runs():
    import __graft_entry__
    __graft_entry__.smoke()
