"""The atan(k / 128) table of so3g_amd/csrc/cr_atan2.h (the device's
correctly rounded atan2 for the CEA pointer): each double-double pair is the
70-digit decimal atan(k / 128) split into its nearest double and the
nearest double of the remainder."""
import os
import re
from decimal import Decimal, localcontext

from helpers import _atan_dec

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   'so3g_amd', 'csrc', 'cr_atan2.h')


def test_atan_table_is_atan_k_over_128():
    text = open(HDR).read()
    body = text[text.index('tab[129][2] = {'):]
    body = body[:body.index('};')]
    pairs = re.findall(r'\{(-?0x[0-9a-fp.+-]+), (-?0x[0-9a-fp.+-]+)\}', body)
    assert len(pairs) == 129
    with localcontext() as ctx:
        ctx.prec = 70
        for k, (hi, lo) in enumerate(pairs):
            v = _atan_dec(Decimal(k) / 128)
            h = float(v)
            assert float.fromhex(hi) == h, k
            assert float.fromhex(lo) == float(v - Decimal(h)), k


def test_pi_constants():
    """pi and pi/2 as double-double pairs in the header."""
    text = open(HDR).read()
    with localcontext() as ctx:
        ctx.prec = 70
        pi = 4 * _atan_dec(Decimal(1))
        for v, exp in ((pi, '+1'), (pi / 2, '+0')):
            h = float(v)
            lo = float(v - Decimal(h))
            assert f'{{{h.hex()}, {lo.hex()}}}'.replace('DD', '') in text.replace('DD', ''), exp
