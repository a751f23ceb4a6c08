"""The JSON log writer (include/clotho_amd/log_writer.hpp) produces the shape
of Boost's write_json of the reference's property trees: nested objects from
dotted keys, every leaf a string, arrays as JSON arrays, one file per record
named <prefix>.<index><suffix>.json."""
import json
import os
import subprocess
import textwrap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = textwrap.dedent(r"""
    #include "clotho_amd/log_writer.hpp"
    #include <vector>
    int main(int argc, char** argv) {
        clotho_amd::json_tree t;
        t.put("performance.runtime", 12345);
        std::vector<double> v = {1.5, 2.0, 0.1};
        t.at("performance").add_value_array("simulate", v.begin(), v.end());
        t.put("stats.sequences_per_allele.mean", 0.25);
        t.put("name", std::string("quote\"d"));
        clotho_amd::json_tree a;
        a.put("position", 0.5);
        a.put("age", 3);
        t.at("allele_space").push_back(a);
        t.at("allele_space").push_back(a);
        clotho_amd::log_writer w(argv[1], ".performance");
        w.write(t);
        w.write(t);
        return 0;
    }
""")


def test_log_writer_shape(tmp_path):
    src = tmp_path / "w.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "w"
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src),
                           "-o", str(exe)])
    prefix = str(tmp_path / "run")
    subprocess.check_call([str(exe), prefix])
    for k in (0, 1):
        d = json.load(open(f"{prefix}.{k}.performance.json"))
        assert d["performance"]["runtime"] == "12345"
        assert [float(x) for x in d["performance"]["simulate"]] == [1.5, 2.0, 0.1]
        assert float(d["stats"]["sequences_per_allele"]["mean"]) == 0.25
        assert d["name"] == 'quote"d'
        assert d["allele_space"] == [{"position": "0.5", "age": "3"}] * 2
