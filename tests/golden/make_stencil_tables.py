"""Writes stencil_tables.json: the velocity numbering tables of the reference's
tests/conftest.py (``_schemes``, lines 19-165), read as text with ast (the
reference is not imported).  Run in the build container only:

    python tests/golden/make_stencil_tables.py /root/reference/tests/conftest.py
"""
import ast
import json
import os
import sys

src = open(sys.argv[1]).read()
tree = ast.parse(src)


class _Ranges(ast.NodeTransformer):
    """list(range(n)) -> [0, ..., n-1] so the table is a plain literal."""

    def visit_Call(self, node):
        if (getattr(node.func, "id", None) == "list" and isinstance(node.args[0], ast.Call)
                and getattr(node.args[0].func, "id", None) == "range"):
            n = ast.literal_eval(node.args[0].args[0])
            return ast.List(elts=[ast.Constant(i) for i in range(n)], ctx=ast.Load())
        return node


tree = _Ranges().visit(tree)
for node in tree.body:
    if isinstance(node, ast.Assign) and getattr(node.targets[0], "id", None) == "_schemes":
        tables = ast.literal_eval(node.value)
        break
else:
    raise SystemExit("_schemes not found")
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stencil_tables.json")
with open(out, "w") as fh:
    json.dump(tables, fh, indent=1)
print("wrote", out, sorted(tables))
