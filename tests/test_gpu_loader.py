"""GPU parity of the text loaders (dss_set_create_text / dss_set_load_file)
against a literal restatement of distribute_strings (src/mpi/readInput.hpp:
40-104): per PE, the same NUL-terminated strings as the MPI rank reads."""
import numpy as np
import pytest

from helpers import file_distributer

pytestmark = pytest.mark.gpu


def _lines(rng, n, lo=0, hi=40, alphabet=(65, 91)):
    return [bytes(rng.integers(alphabet[0], alphabet[1], int(k), dtype=np.uint8)) for k in rng.integers(lo, hi, n)]


def _per_pe(ds):
    chars, off, ln = ds.download()
    pb = ds.pe_begin()
    out = []
    for r in range(ds.p):
        out.append(b"".join(chars[int(off[i]):int(off[i]) + int(ln[i])].tobytes() + b"\0"
                            for i in range(int(pb[r]), int(pb[r + 1]))))
    return out


@pytest.mark.parametrize("p", [1, 3, 4, 16])
@pytest.mark.parametrize("trailing", [True, False])
def test_text_matches_file_distributer(p, trailing):
    from dss_amd import DeviceStringSet
    rng = np.random.default_rng(p * 2 + trailing)
    # bytes 1..255 without '\n': neighbours of 0x0A and bytes with the high bit
    lines = _lines(rng, 700, alphabet=(1, 256))
    lines = [l.replace(b"\n", b"\x0b") for l in lines]
    text = b"\n".join(lines) + (b"\n" if trailing else b"")
    with DeviceStringSet.from_text(text, p) as ds:
        assert _per_pe(ds) == file_distributer(text, p)


def test_text_empty_lines_and_max_size():
    from dss_amd import DeviceStringSet
    rng = np.random.default_rng(9)
    lines = _lines(rng, 400, lo=0, hi=6)
    text = b"\n".join(lines) + b"\n\n\n"
    # cuts that leave a '\n' in every PE's slice: a slice without one is
    # undefined in the reference (shift_left reads past the slice)
    for cut in (0, len(text) // 3, len(text) // 2 + 1, len(text) - 1, len(text)):
        t = text[:cut] if cut else text
        with DeviceStringSet.from_text(text, 4, max_size=cut) as ds:
            assert _per_pe(ds) == file_distributer(t, 4), cut


def test_load_file_equals_text(tmp_path):
    from dss_amd import DeviceStringSet
    rng = np.random.default_rng(4)
    text = b"\n".join(_lines(rng, 20000, 1, 60)) + b"\n"
    path = tmp_path / "in.txt"
    path.write_bytes(text)
    with DeviceStringSet.load_file(str(path), 8) as a, DeviceStringSet.from_text(text, 8) as b:
        assert _per_pe(a) == _per_pe(b) == file_distributer(text, 8)
        # the loaded set goes through the distributed merge sort
        a.merge_sort("IndexedNumStrings", compressed=True)
        chk = a.check()
        assert chk["sorted"] and chk["complete"] and chk["lcp_errors"] == 0


def test_empty_text():
    from dss_amd import DeviceStringSet
    with DeviceStringSet.from_text(b"", 4) as ds:
        assert ds.n == 0
    with DeviceStringSet.from_text(b"\n", 1) as ds:
        assert ds.n == 1 and _per_pe(ds) == [b"\0"]
