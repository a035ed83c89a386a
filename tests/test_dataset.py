"""The dataset's host side on the CPU: the training draw order against the
reference's loop (train.py:364-370), the train/test split
(preprocess_data.py:17-40), and the folder format the loader reads."""
import os
import random

import numpy as np
import pandas as pd

import gsparc_dataset as D
from dataset_fixture import make_dataset


def _reference_draws(data, count, seed):
    """train.py's loop over a list of sample objects, with np.random seeded."""
    np.random.seed(seed)
    train = list(data)
    out = []
    for _ in range(count):
        if not train:
            train = list(data)
            np.random.shuffle(train)
        out.append(train.pop())
    return out


def test_sample_order_is_the_reference_draw_sequence():
    objs = [object() for _ in range(13)]
    ref = _reference_draws(objs, 100, seed=4)
    np.random.seed(4)
    got = D.sample_order(13, 100)
    assert [objs[i] for i in got] == ref
    # the first pass takes the dataset from its end
    assert list(got[:13]) == list(range(12, -1, -1))
    # a seeded RandomState gives the same sequence as the seeded global RNG
    assert np.array_equal(D.sample_order(13, 100, np.random.RandomState(4)), got)


def test_split_dataset(tmp_path):
    root = str(tmp_path)
    make_dataset(root, n=20)
    random.seed(7)
    D.split_dataset(root, ratio=0.8)
    tr = np.loadtxt(os.path.join(root, "train_index.txt"), dtype=str, ndmin=1)
    te = np.loadtxt(os.path.join(root, "test_index.txt"), dtype=str, ndmin=1)
    assert len(tr) == 16 and len(te) == 4
    assert sorted(list(tr) + list(te)) == [f"{i:05d}" for i in range(1, 21)]
    # the split is random.shuffle's permutation of the sorted names
    names = [f"{i:05d}" for i in range(1, 21)]
    random.seed(7)
    random.shuffle(names)
    assert list(tr) == names[:16] and list(te) == names[16:]


def test_dataset_folder_format(tmp_path):
    root = str(tmp_path)
    rx, spectra = make_dataset(root, n=6, H=5, W=8)
    pos, quat = D.read_gateway(root)
    assert np.allclose(pos, [0.5, 1.0, 2.0]) and np.allclose(quat, [0.1, -0.3, 0.2, 0.9])
    assert np.allclose(pd.read_csv(os.path.join(root, "rx_pos.csv")).values, rx, atol=1e-6)
    assert np.array_equal(np.load(os.path.join(root, "spectrum", "00003.npy"), allow_pickle=False), spectra[2])
