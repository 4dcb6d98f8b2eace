"""Extract the frac_coord known-answer vectors of the reference's C unit test
T05_frac_coord (examples/grid/test_recombine.c:378-412) into
t05_frac_coord.json.  The arrays are read from the C source as text; nothing
of the reference is imported or executed.  Run in the container that holds
the reference checkout:  python tests/golden/make_t05_frac_coord.py
"""
import json
import os
import re
import sys

SRC = '/root/reference/examples/grid/test_recombine.c'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 't05_frac_coord.json')


def array(text, name):
    m = re.search(r'double\s+%s\[\]\s*=\s*\{([^}]*)\}' % name, text)
    if not m:
        sys.exit('array %s not found' % name)
    return [float(x) for x in m.group(1).replace('\n', ' ').split(',') if x.strip()]


def main():
    text = open(SRC).read()
    body = text[text.index('int T05_frac_coord()'):]
    body = body[:body.index('return 0;')]
    size = int(re.search(r'grid_size\s*=\s*(\d+)', body).group(1))
    over = int(re.search(r'oversample\s*=\s*(\d+)', body).group(1))
    data = {
        'source': 'examples/grid/test_recombine.c:378-412 (T05_frac_coord)',
        'grid_size': size, 'oversample': over,
        'u': array(body, 'cs'),
        'x': [int(v) for v in array(body, 'ref_x')],
        'fx': [int(v) for v in array(body, 'ref_fx')],
        'note': 'C convention: u in grid cells, x - fx/oversample ~ grid_size/2 + u; '
                'the Python frac_coord(N, Q, p) with p = u / N returns the same (x, fx).',
    }
    assert len(data['u']) == len(data['x']) == len(data['fx']) == 64
    with open(OUT, 'w') as f:
        json.dump(data, f, indent=1)
    print('wrote', OUT)


if __name__ == '__main__':
    main()
