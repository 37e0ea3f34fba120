"""The round-end entry point: __graft_entry__.smoke() on device 0."""

import pytest


@pytest.mark.gpu
def test_graft_entry_smoke(hip):
    import __graft_entry__ as entry

    entry.smoke()
