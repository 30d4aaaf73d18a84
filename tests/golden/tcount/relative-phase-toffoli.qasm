OPENQASM 2.0;
include "qelib1.inc";
// T-count 4 (optimal, up to global phase), 12 gates
qreg q[3];
h q[2];
cx q[2],q[0];
tdg q[0];
t q[2];
cx q[1],q[0];
t q[0];
cx q[2],q[1];
tdg q[1];
cx q[1],q[0];
cx q[2],q[1];
cx q[0],q[2];
h q[2];
