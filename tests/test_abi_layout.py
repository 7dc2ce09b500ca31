"""The drop-in boundary keeps the reference's struct layouts: sizeof,
alignment and every field's offset and size of the 28 structs a caller
compiled against the reference's headers reads or embeds (the embedder
states, the embedded graph and everything nested in them), and the sizes of
the enums and bit-array typedefs among their fields.

tests/golden/abi_layout.json was written by compiling a layout probe against
the reference's own headers (tests/golden/abi_layout.py --reference); this
test compiles the same probe against include/gviz/*.h. CPU only."""
import json
import os
import sys

import pytest

from conftest import GOLDEN

sys.path.insert(0, GOLDEN)
import abi_layout  # noqa: E402


@pytest.fixture(scope="module")
def layouts():
    with open(abi_layout.FIXTURE) as f:
        ref = json.load(f)
    return ref, abi_layout.own_layouts()


@pytest.mark.parametrize("name", [t for t, _ in abi_layout.STRUCTS] + abi_layout.SCALARS)
def test_layout_matches_reference(layouts, name):
    ref, own = layouts
    assert own[name] == ref[name], name


def test_every_state_starts_with_the_embedded_graph(layouts):
    ref, own = layouts
    for t in ("gvizForceEmbedderState", "gvizGRIPState", "gvizTutteState",
              "gvizSpringTutteState"):
        assert own[t]["graph"] == [0, own["gvizEmbeddedGraph"]["sizeof"]]
