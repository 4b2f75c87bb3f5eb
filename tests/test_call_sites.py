"""CPU check that every call into the C ABI passes as many arguments as the
entry point declares: each ``nat.call("mnpbem_...", ...)`` and each
``<lib>.mnpbem_...(...)`` in the package, the tests, the bench and the entry
module is read with ``ast`` and its argument count compared with the ctypes
table (``mnpbem_amd._native.SIGNATURES``, itself checked against the header by
test_abi.py).  Calls with ``*args`` are counted by their fixed parts only
when the tuple is built in the same function."""

import ast
import glob
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _files():
    out = glob.glob(os.path.join(ROOT, "mnpbem_amd", "*.py")) + glob.glob(os.path.join(ROOT, "tests", "*.py"))
    out += [os.path.join(ROOT, "bench.py"), os.path.join(ROOT, "__graft_entry__.py")]
    return sorted(out)


def _tuple_lengths(fn):
    """name -> length of tuples assigned literally in a function body."""
    lens = {}
    for node in ast.walk(fn):
        if isinstance(node, ast.Assign) and isinstance(node.value, ast.Tuple):
            for t in node.targets:
                if isinstance(t, ast.Name) and not any(isinstance(e, ast.Starred) for e in node.value.elts):
                    lens[t.id] = len(node.value.elts)
    return lens


def _count(args, lens):
    n = 0
    for a in args:
        if isinstance(a, ast.Starred):
            if isinstance(a.value, ast.Name) and a.value.id in lens:
                n += lens[a.value.id]
            else:
                return None
        else:
            n += 1
    return n


def test_abi_call_sites_match_signatures():
    from mnpbem_amd._native import SIGNATURES
    bad, seen, done = [], 0, set()
    for path in _files():
        tree = ast.parse(open(path).read(), filename=path)
        scopes = [n for n in ast.walk(tree) if isinstance(n, (ast.FunctionDef, ast.Module))]
        for scope in scopes:
            lens = _tuple_lengths(scope)
            for node in ast.walk(scope):
                if not isinstance(node, ast.Call):
                    continue
                f = node.func
                name, args = None, None
                if isinstance(f, ast.Attribute) and f.attr == "call" and node.args and \
                        isinstance(node.args[0], ast.Constant) and str(node.args[0].value).startswith("mnpbem_"):
                    name, args = node.args[0].value, node.args[1:]
                elif isinstance(f, ast.Attribute) and f.attr.startswith("mnpbem_") and f.attr in SIGNATURES:
                    name, args = f.attr, node.args
                if name is None or (path, node.lineno) in done:
                    continue
                done.add((path, node.lineno))
                assert name in SIGNATURES, (path, node.lineno, name)
                n = _count(args, lens)
                if n is None:
                    continue
                seen += 1
                if n != len(SIGNATURES[name]):
                    bad.append("%s:%d %s: %d arguments, declared %d"
                               % (os.path.relpath(path, ROOT), node.lineno, name, n, len(SIGNATURES[name])))
    assert seen > 50
    assert not bad, "\n".join(bad)
