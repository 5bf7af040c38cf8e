"""Describe the metadata layout of an HDF5 file with h5py: groups and datasets
with their shapes and dtypes, and every attribute's HDF5 class, size, shape
and (for strings) padding and character set.  Used to compare files written
by libosb_io.so with the layout of a reference-held OPS file; runs under any
Python with h5py (``python h5layout.py FILE`` prints JSON)."""
import json
import sys


def _attrs(obj):
    import h5py
    out = {}
    for k in obj.attrs:
        aid = obj.attrs.get_id(k)
        t = aid.get_type()
        d = {"class": type(t).__name__, "size": int(t.get_size()), "shape": list(aid.get_space().shape)}
        if isinstance(t, h5py.h5t.TypeStringID):
            d.update(strpad=int(t.get_strpad()), cset=int(t.get_cset()), vlen=bool(t.is_variable_str()))
        elif isinstance(t, h5py.h5t.TypeIntegerID):
            d.update(signed=int(t.get_sign()), order=int(t.get_order()))
        out[k] = d
    return out


def layout(path):
    import h5py
    res = {}
    with h5py.File(path, "r") as h:
        def visit(name, obj):
            if isinstance(obj, h5py.Group):
                res[name] = {"kind": "group", "attrs": _attrs(obj)}
            else:
                res[name] = {"kind": "dataset", "ndim": obj.ndim, "dtype": obj.dtype.str,
                             "layout": int(obj.id.get_create_plist().get_layout()), "attrs": _attrs(obj)}
        h.visititems(visit)
    return res


def values(path):
    """Attribute values and root one-element datasets (constants), as JSON."""
    import h5py
    res = {"attrs": {}, "consts": {}}
    with h5py.File(path, "r") as h:
        def visit(name, obj):
            res["attrs"][name] = {k: (v.decode() if isinstance(v, bytes) else [int(x) for x in v])
                                  for k, v in obj.attrs.items()}
            if isinstance(obj, h5py.Dataset) and "/" not in name:
                res["consts"][name] = {"value": obj[0].item(), "dtype": obj.dtype.str}
        h.visititems(visit)
    return res


if __name__ == "__main__":
    fn = values if len(sys.argv) > 2 and sys.argv[2] == "values" else layout
    print(json.dumps(fn(sys.argv[1]), sort_keys=True))
