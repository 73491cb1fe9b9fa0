"""Devices, backends, data types and the Buffer record
(reference: runtime/ir/types.py:8-32), with a HIP device / backend in
place of Metal."""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum
from typing import Any, Dict, Optional


class Device(str, Enum):
    CPU = "cpu"
    HIP = "hip"


class Backend(str, Enum):
    CPU = "cpu"
    HIP = "hip"


class DType(str, Enum):
    FR = "fr"
    G1 = "g1"
    G2 = "g2"
    BYTES = "bytes"
    OBJ = "obj"


@dataclass
class Buffer:
    id: str
    device: Device
    dtype: DType
    data: Any
    meta: Optional[Dict[str, Any]] = None
