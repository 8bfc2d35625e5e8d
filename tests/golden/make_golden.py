"""Write the golden fixtures under tests/golden from data files the reference holds.

Nothing here imports or runs reference code; it reads data files as numbers:
* RTCD/Robot/models/panda.h:16-44       -> panda_h_masks.json (the FK mask constants, as numbers)
* RTCD/data/weights/{quad,cubic}/{normal,repeated}/*.txt
                                        -> weights_{kind}_{end}_{nC}_{nT}.txt (least-squares
                                           B-spline fit matrices, row-major nC x nT)
* RTCD/data/benchmark/trajPosesPanda.bin -> traj_poses_panda_256.npy (first 256 fp32 Panda poses)

Usage: python make_golden.py [reference_root]
"""
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    src = open(os.path.join(ref, 'RTCD/Robot/models/panda.h')).read()
    out = {}
    for name in ('sinMask', 'cosMask', 'oneMask', 'lowerBound', 'upperBound', 'baseT'):
        m = re.search(name + r'[^{;]*\{([^}]*)\}', src, re.S)
        out[name] = [float(x.strip().rstrip('f')) for x in m.group(1).split(',') if x.strip()]
    with open(os.path.join(HERE, 'panda_h_masks.json'), 'w') as f:
        json.dump(out, f)
    for nc, nt in ((3, 6), (4, 8), (8, 16), (8, 32), (16, 32)):
        shutil.copy(os.path.join(ref, f'RTCD/data/weights/quad/normal/{nc}_{nt}.txt'),
                    os.path.join(HERE, f'weights_quad_normal_{nc}_{nt}.txt'))
    for kind, end, nc, nt in (('quad', 'repeated', 4, 8), ('quad', 'repeated', 8, 16), ('cubic', 'normal', 4, 16),
                              ('cubic', 'normal', 16, 64), ('cubic', 'repeated', 4, 16), ('cubic', 'repeated', 8, 32)):
        shutil.copy(os.path.join(ref, f'RTCD/data/weights/{kind}/{end}/{nc}_{nt}.txt'),
                    os.path.join(HERE, f'weights_{kind}_{end}_{nc}_{nt}.txt'))
    poses = np.fromfile(os.path.join(ref, 'RTCD/data/benchmark/trajPosesPanda.bin'), dtype=np.float32).reshape(-1, 7)
    np.save(os.path.join(HERE, 'traj_poses_panda_256.npy'), poses[:256])


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
