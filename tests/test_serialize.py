"""Serialization helpers that need no GPU: SEAL's hash (BLAKE2b, RFC 7693) is
pinned against Python's hashlib, including block-boundary lengths and the
digest sizes SEAL uses (32 bytes for parms_id)."""
import hashlib
import os

import pytest

import nexus_amd as nx


@pytest.mark.parametrize("n", [0, 1, 7, 8, 63, 64, 127, 128, 129, 255, 256, 257, 1000, 4096])
@pytest.mark.parametrize("ds", [32, 64, 1, 20])
def test_blake2b_matches_hashlib(n, ds):
    data = os.urandom(n)
    assert nx.blake2b(data, ds) == hashlib.blake2b(data, digest_size=ds).digest()


def test_blake2b_rfc7693_abc():
    # RFC 7693 Appendix A: BLAKE2b-512("abc")
    assert nx.blake2b(b"abc", 64).hex().startswith("ba80a53f981c4d0d6a2797b69f12f6e94c212f14685ac4b74b12bb6fdbffa2d1")
