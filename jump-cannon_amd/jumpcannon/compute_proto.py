"""The graph_compute.Compute wire contract (crates/graph-compute/proto/compute.proto:1-179)
as protobuf message classes, built from a descriptor at import time.

The image has the protobuf runtime but no protoc, so the file descriptor is
assembled here field by field.  Package, message and field names, numbers,
types and labels follow compute.proto exactly: a client generated from that
file (tonic, graph-api) and this module encode and decode the same bytes.
`SERVICE` and `METHODS` name the RPCs for the gRPC method paths.
"""
from __future__ import annotations

from google.protobuf import descriptor_pb2, descriptor_pool, message_factory, struct_pb2

PACKAGE = "graph_compute"
SERVICE = f"{PACKAGE}.Compute"

_F = descriptor_pb2.FieldDescriptorProto
_TYPES = {
    "bytes": _F.TYPE_BYTES, "string": _F.TYPE_STRING, "bool": _F.TYPE_BOOL, "uint32": _F.TYPE_UINT32,
    "uint64": _F.TYPE_UINT64, "float": _F.TYPE_FLOAT,
}

# message -> [(field, number, type or message name, repeated)]
_MESSAGES = {
    "LoadGraphRequest": [("csr", 1, "bytes", False), ("positions", 2, "bytes", False),
                         ("graph_revision", 3, "uint64", False)],
    "LoadGraphResponse": [("ok", 1, "bool", False), ("n_nodes", 2, "uint32", False),
                          ("error", 3, "string", False), ("graph_revision", 4, "uint64", False)],
    "SubscribeRequest": [("graph_id", 1, "string", False), ("layout_id", 2, "string", False),
                         ("params", 3, ".google.protobuf.Struct", False),
                         ("attributes", 4, ".graph_compute.GraphAttributes", False),
                         ("graph_revision", 5, "uint64", False),
                         ("selection_generation", 6, "uint64", False)],
    "GraphAttributes": [("node_class", 1, "bytes", False), ("node_coordination", 2, "bytes", False),
                        ("node_mass", 3, "bytes", False), ("edge_len", 4, "bytes", False)],
    "PositionDelta": [("frame", 1, "uint64", False), ("positions", 2, "bytes", False),
                      ("n_nodes", 3, "uint32", False), ("graph_revision", 4, "uint64", False),
                      ("selection_generation", 5, "uint64", False)],
    "HaloDelta": [("frame", 1, "uint64", False), ("owner_id", 2, "uint32", False),
                  ("node_ids", 3, "bytes", False), ("positions", 4, "bytes", False),
                  ("attributes", 5, ".graph_compute.GraphAttributes", False)],
    "EngineDescriptor": [("id", 1, "string", False), ("display_name", 2, "string", False),
                         ("description", 3, "string", False), ("kind", 4, "string", False)],
    "ListEnginesRequest": [],
    "ListEnginesResponse": [("engines", 1, ".graph_compute.EngineDescriptor", True),
                            ("default_id", 2, "string", False)],
    "HealthRequest": [],
    "HealthResponse": [("ok", 1, "bool", False), ("n_nodes", 2, "uint32", False),
                       ("frame", 3, "uint64", False), ("requested_layout_id", 4, "string", False),
                       ("effective_layout_id", 5, "string", False), ("fallback_reason", 6, "string", False),
                       ("graph_revision", 7, "uint64", False), ("selection_generation", 8, "uint64", False)],
    "CoarsenSettings": [("max_levels", 1, "uint32", False), ("target_size", 2, "uint32", False),
                        ("gt_max_hops", 3, "uint32", False), ("w_proximity", 4, "float", False),
                        ("w_size", 5, "float", False), ("w_connection", 6, "float", False),
                        ("w_neighborhood", 7, "float", False), ("w_degree", 8, "float", False)],
    "FocusRequest": [("graph_id", 1, "string", False), ("focal_node", 2, "uint32", False),
                     ("capacities", 3, "uint32", True), ("alpha", 4, "float", False),
                     ("coarsen", 5, ".graph_compute.CoarsenSettings", False)],
    "HybridFrame": [("n_nodes", 1, "uint32", False), ("node_refs", 2, "bytes", False),
                    ("node_levels", 3, "bytes", False), ("positions", 4, "bytes", False),
                    ("edges", 5, "bytes", False), ("edge_levels", 6, "bytes", False),
                    ("n_edges", 7, "uint32", False)],
}

# rpc -> (request, response, client streams, server streams)
METHODS = {
    "Subscribe": ("SubscribeRequest", "PositionDelta", False, True),
    "Health": ("HealthRequest", "HealthResponse", False, False),
    "ListEngines": ("ListEnginesRequest", "ListEnginesResponse", False, False),
    "TopoFisheye": ("FocusRequest", "HybridFrame", True, True),
    "ExchangeHalo": ("HaloDelta", "HaloDelta", True, True),
    "LoadGraph": ("LoadGraphRequest", "LoadGraphResponse", False, False),
}


def _file() -> descriptor_pb2.FileDescriptorProto:
    f = descriptor_pb2.FileDescriptorProto(name="graph_compute/compute.proto", package=PACKAGE,
                                           syntax="proto3", dependency=["google/protobuf/struct.proto"])
    for name, fields in _MESSAGES.items():
        m = f.message_type.add(name=name)
        for fname, num, typ, repeated in fields:
            fd = m.field.add(name=fname, number=num, json_name=_camel(fname),
                             label=_F.LABEL_REPEATED if repeated else _F.LABEL_OPTIONAL)
            if typ.startswith("."):
                fd.type = _F.TYPE_MESSAGE
                fd.type_name = typ
            else:
                fd.type = _TYPES[typ]
    svc = f.service.add(name="Compute")
    for rpc, (req, resp, cs, ss) in METHODS.items():
        svc.method.add(name=rpc, input_type=f".{PACKAGE}.{req}", output_type=f".{PACKAGE}.{resp}",
                       client_streaming=cs, server_streaming=ss)
    return f


def _camel(name: str) -> str:
    head, *rest = name.split("_")
    return head + "".join(w[:1].upper() + w[1:] for w in rest)


_pool = descriptor_pool.DescriptorPool()
_pool.AddSerializedFile(struct_pb2.DESCRIPTOR.serialized_pb)
_pool.AddSerializedFile(_file().SerializeToString())

FILE = _pool.FindFileByName("graph_compute/compute.proto")
Struct = message_factory.GetMessageClass(_pool.FindMessageTypeByName("google.protobuf.Struct"))
globals().update({name: message_factory.GetMessageClass(_pool.FindMessageTypeByName(f"{PACKAGE}.{name}"))
                  for name in _MESSAGES})


def message(name: str):
    """The message class of compute.proto's `name`."""
    return globals()[name]


def method_path(rpc: str) -> str:
    return f"/{SERVICE}/{rpc}"
