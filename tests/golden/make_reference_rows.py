"""Extracts the measured and expected values the reference's own test fixture holds
(lib/pathtracer/test/fixtures/run-0.77.0-pascal.json: the user's 0.77.0 run on an NVIDIA Pascal card) into
reference_rows.json: data only -- each row's case name, measured mean luminance, relative standard error and
expected value. Run in a checkout that has the reference at /root/reference."""
import json
import os
import sys

SRC = '/root/reference/lib/pathtracer/test/fixtures/run-0.77.0-pascal.json'
KEEP = ['furnace', 'haze', 'photometry', 'illuminance', 'cosine-law', 'bsdf-parity', 'optics-refraction', 'optics-farfield',
        'optics-balance']


def main():
    d = json.load(open(sys.argv[1] if len(sys.argv) > 1 else SRC))
    out = {'source': 'lib/pathtracer/test/fixtures/run-0.77.0-pascal.json', 'tests': {}}
    for k in KEEP:
        t = d['tests'][k]
        rows = [{f: r.get(f) for f in ('case', 'meanLuma', 'relSE', 'expected', 'value', 'measured', 'rays') if f in r}
                for r in t.get('rows', [])]
        out['tests'][k] = {'pixels': t.get('pixels'), 'rows': rows}
    here = os.path.dirname(os.path.abspath(__file__))
    json.dump(out, open(os.path.join(here, 'reference_rows.json'), 'w'), indent=1)


if __name__ == '__main__':
    main()
