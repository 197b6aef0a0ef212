"""Philox4x32-10 of the oracle against published known-answer vectors."""
from oracle import binding as ob

# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key) -> out
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_kat_vectors():
    for ctr, key, out in KAT:
        assert tuple(ob.philox(ctr, key)) == out


def test_u53_range():
    L = ob.lib()
    assert L.or_u53(0, 0) == 0.0
    assert L.or_u53(0xffffffff, 0xffffffff) == 1.0 - 2.0 ** -53
    assert L.or_u53_open0(0, 0) == 2.0 ** -53
    assert L.or_u53_open0(0xffffffff, 0xffffffff) == 1.0
