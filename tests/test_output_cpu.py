"""Host-side output formats (VTKWriter / StatsModule restatements): java.util.Formatter's %g,
Writer.splitFileName and the VTK binary data-array encoding, against known answers."""
import base64
import struct

import pytest

from starfish_amd.output import _data_array, java_format_g, split_file_name


@pytest.mark.parametrize("value,prec,text", [
    (1.0, 6, "1.00000"),            # Java keeps trailing zeros (C's %g prints 1)
    (123456.0, 6, "123456"),
    (1234567.0, 6, "1.23457e+06"),
    (0.0001, 6, "0.000100000"),     # 1e-4 is still plain notation
    (1e-05, 6, "1.00000e-05"),
    (0.5, 4, "0.5000"),
    (1e10, 4, "1.000e+10"),
    (-2.5e-07, 4, "-2.500e-07"),
    (12345.0, 4, "1.235e+04"),      # HALF_UP on the decimal digits (C rounds the tie to 1.234e+04)
    (100.0, 4, "100.0"),
    (1000.0, 4, "1000"),
    (9.99995e-05, 4, "0.0001000"),  # the notation follows the rounded value
    (0.0, 6, "0.00000"),
    (0.0, 4, "0.000"),
    (float("nan"), 4, "NaN"),
    (float("inf"), 6, "Infinity"),
])
def test_java_general_format(value, prec, text):
    assert java_format_g(value, prec) == text


def test_split_file_name():
    assert split_file_name("parts.vtp") == ["parts", ".vtp", "", "parts"]
    assert split_file_name("out/parts.vtp") == ["out/parts", ".vtp", "out", "parts"]
    # no extension: the reference splits before the last character
    assert split_file_name("parts") == ["part", "s", "", "part"]


def test_binary_data_array_layout():
    lines = []
    _data_array(lines, "v", [1.0, 2.0, -3.5], 3, "binary")
    assert lines[0] == '<DataArray Name="v" type="Float64" NumberOfComponents="3" format="binary">'
    raw = base64.b64decode(lines[1])
    # int32 byte count, the doubles, then the 4 bytes ByteBuffer.allocate(n + Long.BYTES) leaves zero
    assert raw == struct.pack("<i3d", 24, 1.0, 2.0, -3.5) + b"\0" * 4
    assert lines[2] == "</DataArray>"
    lines = []
    _data_array(lines, "k", [3, -1], 1, "binary")
    assert base64.b64decode(lines[1]) == struct.pack("<i2i", 8, 3, -1) + b"\0" * 4
    assert 'type="Int32"' in lines[0]


def test_ascii_data_array_layout():
    lines = []
    _data_array(lines, "r", [0.5, 0.0], 1, "ascii")
    assert lines == ['<DataArray Name="r" type="Float64" NumberOfComponents="1" format="ascii">',
                     "0.500000 0.00000 \n</DataArray>"]
