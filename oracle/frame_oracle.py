"""CPU ORACLE — TEST INFRASTRUCTURE ONLY (video frame message).

Restates serverSendFrame without libjpeg (lib/isaac/isaac_communicator.hpp:217-287,
ISAAC_JPEG off): {"payload": "data:image/raw-rgba;base64,<payload>"} where the
payload is boost's base64_from_binary<transform_width<const unsigned char*, 6, 8>>
over the frame bytes (isaac_communicator.hpp:251-263).  Boost is not in the
reference; its iterators emit one character per 6 bits, most significant first,
the last partial group zero-filled, and add no '=' padding: that is RFC 4648
base64 (Python's base64 module, a library routine) with the '=' removed.

Only tests/ may import this module; the product never does.
"""
from __future__ import annotations

import base64


def b64_unpadded(data: bytes) -> bytes:
    return base64.b64encode(bytes(data)).rstrip(b"=")


def frame_message(rgba: bytes) -> bytes:
    return b'{"payload": "data:image/raw-rgba;base64,' + b64_unpadded(rgba) + b'"}'
