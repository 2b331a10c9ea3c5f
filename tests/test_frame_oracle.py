"""The frame message restatement (serverSendFrame without libjpeg) on known
answers: base64 of "Man" / "Ma" / "M" is "TWFu" / "TWE" / "TQ" (RFC 4648
section 4 examples with the '=' padding the reference's boost iterators do not
emit), the character count is ceil(8n/6), and the message framing."""
from oracle import frame_oracle as F


def test_known_answers():
    assert F.b64_unpadded(b"Man") == b"TWFu"
    assert F.b64_unpadded(b"Ma") == b"TWE"
    assert F.b64_unpadded(b"M") == b"TQ"
    assert F.b64_unpadded(b"") == b""
    assert F.b64_unpadded(bytes([0xFF, 0xEF, 0xBE])) == b"/+++"
    assert F.b64_unpadded(b"foobar") == b"Zm9vYmFy"
    assert F.b64_unpadded(b"fooba") == b"Zm9vYmE"


def test_lengths():
    for n in range(0, 40):
        assert len(F.b64_unpadded(bytes(range(n)))) == (8 * n + 5) // 6


def test_message_framing():
    msg = F.frame_message(bytes([1, 2, 3, 4]))
    assert msg == b'{"payload": "data:image/raw-rgba;base64,AQIDBA"}'
