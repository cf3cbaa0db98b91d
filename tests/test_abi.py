"""CPU checks of the C-ABI boundary: the library loads, exports every entry
point include/nexus_ckks.h declares, and fails loudly (no CPU fallback) when no
GPU is present."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_symbols():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(nx_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) > 40
    assert "nx_relinearize" in syms and "nx_rescale_to_next" in syms and "nx_rotate_vector" in syms


def test_library_exports_every_declared_symbol():
    import nexus_amd

    lib = ctypes.CDLL(nexus_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_cpu_fallback_without_gpu():
    import nexus_amd as nx

    if nx.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(nx.NexusError) as e:
        nx.Context(12, [58, 40, 58])
    assert e.value.code == -4  # NX_ERR_DEVICE


def test_error_codes_and_messages():
    import nexus_amd as nx

    with pytest.raises(nx.NexusError) as e:
        nx.Context(12, [58])  # needs a data prime and a special prime
    assert e.value.code == -1
    assert "special prime" in str(e.value)


def test_coeff_modulus_create_host():
    """CoeffModulus::Create runs on the host: equal to the oracle's restatement
    (pinned by SEAL's KATs); bit sizes above SEAL's 60-bit cap
    (util/defines.h:40) are rejected, since the lazy 128-bit accumulators of the
    inner products assume residues below 2^60."""
    import nexus_amd as nx
    import oracle as orc

    for n, bits in ((1 << 13, [60, 40, 40, 60]), (1 << 16, [58] + [40] * 18 + [58]), (1 << 12, [30, 45, 51, 2 + 58])):
        assert nx.coeff_modulus_create(n, bits) == orc.coeff_modulus_create(n, bits)
    for bad in ([61, 40, 61], [40, 1], [14, 40]):  # 14 bits: no prime = 1 mod 2^14 below 2^14
        with pytest.raises(nx.NexusError) as e:
            nx.coeff_modulus_create(1 << 13, bad)
        assert e.value.code in (-1, -2)
