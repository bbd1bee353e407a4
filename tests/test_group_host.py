"""Host-side checks of the group API that need no GPU."""
import pytest

from powder6_amd import problems as PR
from powder6_amd import pt


def test_group_without_device_names_member_and_fails_loudly():
    pbs = [PR.uneven_pattern(i, chains=4) for i in (0, 1)]
    with pytest.raises(pt.PTError, match=r'group member 0: MI355X HIP device is unavailable'):
        pt.Group(pbs)


def test_group_needs_problems():
    with pytest.raises(pt.PTError, match='at least one problem'):
        pt.Group([])
