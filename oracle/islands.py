"""Island order of dWorldQuickStep (TEST INFRASTRUCTURE ONLY).

Restates BuildIslandsAndEstimateStepperMemoryRequirements (ode/src/util.cpp:
724-855) over ODE's list structure: bodies head-inserted into the world list
(addObjectToList, ode.cpp:54-60), joints head-inserted into each attached
body's joint-node list by dJointAttach (ode.cpp:1383-1439).  Returns the
bodies and joints in the order the island stepper sees them.
"""


def island_order(n_bodies, joints, enabled=None):
    """n_bodies bodies created in order 0..n-1; joints: list of (b1, b2) in
    attach order with b2 = -1 for the world (b1 >= 0).  Returns (bodies,
    joints) in stepping order (joint indices refer to the input list)."""
    enabled = enabled if enabled is not None else [True] * n_bodies
    world = list(range(n_bodies - 1, -1, -1))  # head insertion
    nodes = {b: [] for b in range(n_bodies)}   # per body: (joint, other body), most recent first
    for k, (b1, b2) in enumerate(joints):
        nodes[b1].insert(0, (k, b2))
        if b2 >= 0:
            nodes[b2].insert(0, (k, b1))
    btag = [0] * n_bodies
    jtag = [False] * len(joints)
    out_b, out_j = [], []
    for bb in world:
        if btag[bb] or not enabled[bb]:
            continue
        btag[bb] = 1
        out_b.append(bb)
        stack = []
        b = bb
        while True:
            for k, other in nodes[b]:
                if jtag[k]:
                    continue
                jtag[k] = True
                out_j.append(k)
                if other >= 0 and btag[other] <= 0:
                    btag[other] = 1
                    stack.append(other)
            if not stack:
                break
            b = stack.pop()
            out_b.append(b)
    return out_b, out_j
