"""The .cube metadata schema (wavelet_compress.proto:7-174) rebuilt as a
protobuf descriptor at run time, so the Python protobuf runtime can parse and
re-serialise what the oracle and the product write (there is no protoc here).
Test infrastructure only.
"""
from google.protobuf import descriptor_pb2, descriptor_pool, message_factory

F = descriptor_pb2.FieldDescriptorProto

_cache = {}


def _field(msg, name, number, ftype, label=F.LABEL_OPTIONAL, type_name=None,
           packed=False, default=None):
    f = msg.field.add()
    f.name, f.number, f.type, f.label = name, number, ftype, label
    if type_name:
        f.type_name = type_name
    if packed:
        f.options.packed = True
    if default is not None:
        f.default_value = default
    return f


def messages():
    if _cache:
        return _cache
    fd = descriptor_pb2.FileDescriptorProto()
    fd.name = "wavelet_compress_rebuilt.proto"
    fd.package = "wvtest"
    fd.syntax = "proto2"

    cb = fd.message_type.add()
    cb.name = "CubeletBuffer"
    dt = cb.enum_type.add()
    dt.name = "DataType"
    for n, v in (("UINT8", 1), ("INT32", 2), ("FLOAT32", 3)):
        e = dt.value.add(); e.name, e.number = n, v
    ca = cb.enum_type.add()
    ca.name = "CompressionAlgorithm"
    for n, v in (("NONE", 0), ("ZLIB", 1), ("WAVELET", 2)):
        e = ca.value.add(); e.name, e.number = n, v
    R = F.LABEL_REQUIRED
    _field(cb, "width", 1, F.TYPE_UINT32, R)
    _field(cb, "height", 2, F.TYPE_UINT32, R)
    _field(cb, "depth", 3, F.TYPE_UINT32, R)
    _field(cb, "x_offset", 4, F.TYPE_UINT32, default="0")
    _field(cb, "y_offset", 5, F.TYPE_UINT32, default="0")
    _field(cb, "z_offset", 6, F.TYPE_UINT32, default="0")
    _field(cb, "data_type", 7, F.TYPE_ENUM, R, ".wvtest.CubeletBuffer.DataType")
    _field(cb, "maximum_value", 15, F.TYPE_INT32, default="0")
    _field(cb, "byte_count", 8, F.TYPE_UINT32, R)
    _field(cb, "compression_algorithm", 9, F.TYPE_ENUM,
           type_name=".wvtest.CubeletBuffer.CompressionAlgorithm", default="NONE")
    _field(cb, "compression_parameters", 14, F.TYPE_MESSAGE,
           type_name=".wvtest.CompressionParametersBuffer")
    _field(cb, "checksum", 10, F.TYPE_UINT32)
    _field(cb, "data_file_name", 11, F.TYPE_STRING)
    _field(cb, "data_file_index", 12, F.TYPE_UINT32)
    _field(cb, "data_file_offset", 13, F.TYPE_UINT64, default="0")

    cp = fd.message_type.add()
    cp.name = "CompressionParametersBuffer"
    wa = cp.enum_type.add()
    wa.name = "WaveletAlgorithm"
    for n, v in (("HAAR", 1), ("CDF97", 2), ("CDF53", 3)):
        e = wa.value.add(); e.name, e.number = n, v
    qa = cp.enum_type.add()
    qa.name = "QuantizationAlgorithm"
    for n, v in (("LOG", 1), ("UNIFORM", 2), ("LLOYD", 3)):
        e = qa.value.add(); e.name, e.number = n, v
    _field(cp, "original_width", 14, F.TYPE_UINT32)
    _field(cp, "original_height", 15, F.TYPE_UINT32)
    _field(cp, "original_depth", 16, F.TYPE_UINT32)
    _field(cp, "original_data_type", 17, F.TYPE_ENUM,
           type_name=".wvtest.CubeletBuffer.DataType")
    _field(cp, "wavelet_transform_steps_x", 1, F.TYPE_INT32, R)
    _field(cp, "wavelet_transform_steps_y", 2, F.TYPE_INT32, R)
    _field(cp, "wavelet_transform_steps_z", 3, F.TYPE_INT32, R)
    _field(cp, "wavelet_algorithm", 4, F.TYPE_ENUM, R,
           ".wvtest.CompressionParametersBuffer.WaveletAlgorithm", default="CDF97")
    _field(cp, "transform_2d", 18, F.TYPE_BOOL, default="false")
    _field(cp, "threshold_fraction", 6, F.TYPE_FLOAT)
    _field(cp, "threshold_value", 7, F.TYPE_FLOAT)
    _field(cp, "quant_bin_count", 8, F.TYPE_INT32, R)
    _field(cp, "quantization_algorithm", 9, F.TYPE_ENUM,
           type_name=".wvtest.CompressionParametersBuffer.QuantizationAlgorithm",
           default="LOG")
    _field(cp, "quant_max_value", 10, F.TYPE_FLOAT)
    _field(cp, "quant_bin_boundaries", 11, F.TYPE_FLOAT, F.LABEL_REPEATED, packed=True)
    _field(cp, "quant_bin_values", 12, F.TYPE_FLOAT, F.LABEL_REPEATED, packed=True)
    _field(cp, "huffman_decode_table", 13, F.TYPE_INT32, F.LABEL_REPEATED, packed=True)
    _field(cp, "huffman_dup_key", 19, F.TYPE_INT32)
    _field(cp, "huffman_dup_value", 20, F.TYPE_INT32)

    ix = fd.message_type.add()
    ix.name = "CubeletIndexBuffer"
    _field(ix, "width", 1, F.TYPE_UINT32)
    _field(ix, "height", 2, F.TYPE_UINT32)
    _field(ix, "depth", 3, F.TYPE_UINT32)
    _field(ix, "data_files", 4, F.TYPE_STRING, F.LABEL_REPEATED)
    _field(ix, "cubelets", 5, F.TYPE_MESSAGE, F.LABEL_REPEATED, ".wvtest.CubeletBuffer")

    pool = descriptor_pool.DescriptorPool()
    pool.Add(fd)
    for name in ("CubeletBuffer", "CompressionParametersBuffer", "CubeletIndexBuffer"):
        desc = pool.FindMessageTypeByName("wvtest." + name)
        _cache[name] = message_factory.GetMessageClass(desc)
    return _cache


def parse_cube_file(data):
    """Splits a .cube byte string into (header, [(CubeletBuffer, raw meta
    bytes, payload)], footer index or None, raw index bytes)."""
    import struct
    m = messages()
    assert data[:17] == b"SCU cubelets 1.0\n"
    header = data[:64]
    pos = 64
    cubes = []
    index = None
    raw_index = None
    while pos < len(data):
        (n,) = struct.unpack_from("<I", data, pos)
        pos += 4
        if n == 0xFFFFFFFF:
            (ilen,) = struct.unpack_from("<i", data, len(data) - 8)
            assert data[-4:] == b"cube"
            raw_index = data[pos:pos + ilen]
            assert pos + ilen == len(data) - 8
            index = m["CubeletIndexBuffer"]()
            index.ParseFromString(raw_index)
            break
        raw = data[pos:pos + n]
        cb = m["CubeletBuffer"]()
        cb.ParseFromString(raw)
        pos += n
        payload = data[pos:pos + cb.byte_count]
        pos += cb.byte_count
        cubes.append((cb, raw, payload))
    return header, cubes, index, raw_index
