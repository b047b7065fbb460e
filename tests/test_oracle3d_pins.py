"""Pins of the 3D offline_log compressor's oracle against the reference's
header-only helpers (compiled into oracle/_ref) and against the mathematics of
the per-tetrahedron and per-vertex bounds; see tests/_pins3d.py."""
import pytest

from _libs import Reference, have_reference, load_oracle
from _pins3d import PINS

pytestmark = pytest.mark.skipif(not have_reference(), reason="reference build absent (oracle/_ref)")


@pytest.mark.parametrize("pin", PINS, ids=lambda f: f.__name__)
def test_pin(pin):
    pin(load_oracle(), Reference())
