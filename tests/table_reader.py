"""Read the generated C coefficient tables (boundary_tables.h) into Python lists."""
import ast
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = os.path.join(HERE, "..", "lbm3_amd", "csrc", "generated", "boundary_tables.h")


def load_tables(path=TABLES):
    text = open(path).read()
    out = {}
    for m in re.finditer(r"LBM3_BT_CONST [a-z ]+ LBM3_BT_(\w+)((?:\[\d+\])+) = (.*?);\n", text):
        out[m.group(1)] = ast.literal_eval(m.group(3).replace("{", "[").replace("}", "]"))
    return out
