"""Write matching_networkx.json: maximum-weight matchings of random graphs
computed with networkx.max_weight_matching (networkx 3.x, the published
Galil/Edmonds primal-dual implementation), the reference answers for the
blossom solver in libqcf.

Each case: n vertices, edge list (u, v, w) with integer weights, and
networkx's matching weight and size (maxcardinality=False, the default).
"""
import json
import os

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(2024)
    cases = []
    for k in range(120):
        n = int(rng.integers(1, 19))
        dens = float(rng.choice([0.3, 0.6, 1.0]))
        wmax = int(rng.choice([3, 20, 1000, 10**6]))
        G = nx.Graph()
        G.add_nodes_from(range(n))
        edges = []
        for u in range(n):
            for v in range(u + 1, n):
                if rng.random() < dens:
                    w = int(rng.integers(1, wmax + 1))
                    edges.append((u, v, w))
                    G.add_edge(u, v, weight=w)
        m = nx.max_weight_matching(G)
        weight = sum(G[u][v]["weight"] for u, v in m)
        cases.append({"n": n, "edges": edges, "weight": int(weight), "size": len(m)})
    with open(os.path.join(HERE, "matching_networkx.json"), "w") as f:
        json.dump({"generator": f"networkx {nx.__version__} max_weight_matching", "cases": cases}, f)


if __name__ == "__main__":
    main()
