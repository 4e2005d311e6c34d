"""The ``.trees`` file format: a kastore key-value store of typed arrays
(c/subprojects/kastore/kastore.h:108-166, kastore.c:154-340) holding the
table columns under the keys of c/tskit/tables.c (nodes 2551-2595, edges
3244-3301, sites 4000-4041, mutations 4803-4840, collection 11588-11700,
indexes 11727-11753).

Text columns are u8 arrays with u64 offsets ("<name>_offset"); collection
properties (format name, uuid, time units, metadata) are i8.

Layout: a 64-byte header (magic, version 1.x as two u16, item count u32, file
size u64), one 64-byte descriptor per item (type u8; key start, key length,
array start, array length as u64 at bytes 8, 16, 24, 32), the keys packed
back to back, then the arrays, each aligned to 8 bytes.  Items are stored in
key order (readers look keys up by binary search).

``load`` reads the node, edge, site and mutation columns (other tables,
metadata and provenance are not on the statistics path and are skipped);
``dump`` writes a file the reference's loader accepts, with empty
individual, population, migration and provenance tables.  Nothing in a file
is executed: arrays are read with ``numpy.frombuffer``.
"""
import uuid as _uuid

import numpy as np

MAGIC = b"\x89KAS\r\n\x1a\n"
HEADER = 64
DESCRIPTOR = 64
ALIGN = 8
# kastore.h:108-117
TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64,
         np.float32, np.float64]
FORMAT_NAME = b"tskit.trees"
FORMAT_VERSION = (12, 7)  # core.h:206-209


class FileFormatError(Exception):
    pass


def read(path):
    """All items of a kastore file: {key (str): numpy array}."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < HEADER or data[:8] != MAGIC:
        raise FileFormatError(f"{path}: not a kastore file")
    major = int(np.frombuffer(data, np.uint16, 1, 8)[0])
    if major != 1:
        raise FileFormatError(f"{path}: kastore format version {major} not supported")
    num_items = int(np.frombuffer(data, np.uint32, 1, 12)[0])
    file_size = int(np.frombuffer(data, np.uint64, 1, 16)[0])
    if file_size != len(data) or HEADER + num_items * DESCRIPTOR > file_size:
        raise FileFormatError(f"{path}: truncated or malformed kastore file")
    items = {}
    for j in range(num_items):
        d = HEADER + j * DESCRIPTOR
        kind = data[d]
        if kind >= len(TYPES):
            raise FileFormatError(f"{path}: bad item type {kind}")
        key_start, key_len, arr_start, arr_len = (
            int(x) for x in np.frombuffer(data, np.uint64, 4, d + 8))
        dt = np.dtype(TYPES[kind])
        if key_start + key_len > file_size or arr_start + arr_len * dt.itemsize > file_size:
            raise FileFormatError(f"{path}: item {j} out of bounds")
        key = data[key_start:key_start + key_len].decode("utf-8")
        items[key] = np.frombuffer(data, dt, arr_len, arr_start).copy()
    return items


def write(path, items):
    """Writes {key: numpy array} as a kastore file, keys in sorted order."""
    keys = sorted(items, key=lambda k: k.encode("utf-8"))
    enc = [k.encode("utf-8") for k in keys]
    arrays = []
    for k in keys:
        a = np.ascontiguousarray(items[k])
        if a.dtype.type not in TYPES:
            raise TypeError(f"{k}: unsupported dtype {a.dtype}")
        arrays.append(a)
    offset = HEADER + len(keys) * DESCRIPTOR
    key_starts = []
    for e in enc:
        key_starts.append(offset)
        offset += len(e)
    arr_starts = []
    for a in arrays:
        offset += (-offset) % ALIGN
        arr_starts.append(offset)
        offset += a.nbytes
    buf = bytearray(offset)
    buf[:8] = MAGIC
    buf[8:12] = np.array([1, 0], np.uint16).tobytes()
    buf[12:16] = np.array([len(keys)], np.uint32).tobytes()
    buf[16:24] = np.array([offset], np.uint64).tobytes()
    for j, (e, a) in enumerate(zip(enc, arrays)):
        d = HEADER + j * DESCRIPTOR
        buf[d] = TYPES.index(a.dtype.type)
        buf[d + 8:d + 40] = np.array([key_starts[j], len(e), arr_starts[j], a.size],
                                     np.uint64).tobytes()
        buf[key_starts[j]:key_starts[j] + len(e)] = e
        buf[arr_starts[j]:arr_starts[j] + a.nbytes] = a.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(buf))


def load(path):
    """A TableCollection from a ``.trees`` file (nodes, edges, sites,
    mutations; time units)."""
    from .tables import TableCollection

    it = read(path)
    name = it.get("format/name")
    if name is None or name.tobytes() != FORMAT_NAME:
        raise FileFormatError(f"{path}: not a tskit.trees file")
    tc = TableCollection(float(it["sequence_length"][0]))
    if "time_units" in it:
        tc.time_units = it["time_units"].tobytes().decode("utf-8")
    tc.set_nodes(it["nodes/flags"].astype(np.uint32), it["nodes/time"].astype(np.float64))
    tc.set_edges(it["edges/left"], it["edges/right"], it["edges/parent"].astype(np.int32),
                 it["edges/child"].astype(np.int32))
    tc.site_position = it["sites/position"].astype(np.float64)
    tc.ancestral_state = it["sites/ancestral_state"].view(np.int8).copy()
    tc.ancestral_state_offset = it["sites/ancestral_state_offset"].astype(np.uint64)
    tc.mutation_site = it["mutations/site"].astype(np.int32)
    tc.mutation_node = it["mutations/node"].astype(np.int32)
    tc.mutation_parent = it["mutations/parent"].astype(np.int32)
    tc.derived_state = it["mutations/derived_state"].view(np.int8).copy()
    tc.derived_state_offset = it["mutations/derived_state_offset"].astype(np.uint64)
    return tc


def dump(tc, path):
    """Writes a TableCollection as a ``.trees`` file the reference loads."""
    N, E, S, Mu = tc.num_nodes, tc.num_edges, tc.num_sites, tc.num_mutations
    ids = lambda n: np.full(n, -1, np.int32)  # noqa: E731
    empty = lambda n: (np.zeros(0, np.uint8), np.zeros(n + 1, np.uint64))  # noqa: E731
    items = {
        "format/name": np.frombuffer(FORMAT_NAME, np.int8),
        "format/version": np.array(FORMAT_VERSION, np.uint32),
        "sequence_length": np.array([tc.sequence_length], np.float64),
        "uuid": np.frombuffer(str(_uuid.uuid4()).encode(), np.int8),
        "time_units": np.frombuffer(tc.time_units.encode(), np.int8),
        "metadata": np.zeros(0, np.int8),
        "metadata_schema": np.zeros(0, np.int8),
        "nodes/time": np.asarray(tc.node_time, np.float64),
        "nodes/flags": np.asarray(tc.node_flags, np.uint32),
        "nodes/population": ids(N),
        "nodes/individual": ids(N),
        "edges/left": np.asarray(tc.edge_left, np.float64),
        "edges/right": np.asarray(tc.edge_right, np.float64),
        "edges/parent": np.asarray(tc.edge_parent, np.int32),
        "edges/child": np.asarray(tc.edge_child, np.int32),
        "sites/position": np.asarray(tc.site_position, np.float64),
        "sites/ancestral_state": np.asarray(tc.ancestral_state).view(np.uint8),
        "sites/ancestral_state_offset": np.asarray(tc.ancestral_state_offset, np.uint64),
        "mutations/site": np.asarray(tc.mutation_site, np.int32),
        "mutations/node": np.asarray(tc.mutation_node, np.int32),
        "mutations/parent": np.asarray(tc.mutation_parent, np.int32),
        "mutations/derived_state": np.asarray(tc.derived_state).view(np.uint8),
        "mutations/derived_state_offset": np.asarray(tc.derived_state_offset, np.uint64),
        "individuals/flags": np.zeros(0, np.uint32),
        "migrations/left": np.zeros(0, np.float64),
        "migrations/right": np.zeros(0, np.float64),
        "migrations/node": ids(0),
        "migrations/source": ids(0),
        "migrations/dest": ids(0),
        "migrations/time": np.zeros(0, np.float64),
    }
    ragged = {"nodes/metadata": N, "edges/metadata": E, "sites/metadata": S,
              "mutations/metadata": Mu, "individuals/location": 0, "individuals/parents": 0,
              "individuals/metadata": 0, "populations/metadata": 0, "migrations/metadata": 0,
              "provenances/timestamp": 0, "provenances/record": 0}
    for key, n in ragged.items():
        data, off = empty(n)
        if key == "individuals/location":
            data = np.zeros(0, np.float64)
        elif key == "individuals/parents":
            data = ids(0)
        items[key] = data
        items[key + "_offset"] = off
    write(path, items)
