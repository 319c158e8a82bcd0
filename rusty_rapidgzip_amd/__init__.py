"""MI355X-native gzip/DEFLATE decoder with rusty_rapidgzip's ``read_gz`` surface."""
from .api import (  # noqa: F401
    Config,
    DecodeStats,
    Decoder,
    DeflateError,
    DeviceError,
    GzipError,
    IoError,
    Profile,
    RapidgzipError,
    SinkClosed,
    Verbosity,
    decode_all,
    decode_bytes,
    read_gz,
)

__all__ = [
    "Config",
    "DecodeStats",
    "Decoder",
    "DeflateError",
    "DeviceError",
    "GzipError",
    "IoError",
    "Profile",
    "RapidgzipError",
    "SinkClosed",
    "Verbosity",
    "decode_all",
    "decode_bytes",
    "read_gz",
]
